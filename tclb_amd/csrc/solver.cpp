// solver.cpp -- see solver.h.  Handler semantics follow src/Handlers/
// (file named at each class); scheduling follows vHandler::Now/Next
// (src/Handlers/vHandler.h:68-93) and acSolve::Init (src/Handlers/acSolve.cpp).
#include "solver.h"

#include <sys/stat.h>
#include <zlib.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <fstream>
#include <set>
#include <sstream>
#include <stdexcept>

namespace tclbx {

namespace {

[[noreturn]] void fail(const std::string& m) { throw std::runtime_error(m); }

int myround(double v) { return v > 0 ? (int)(v + 0.5) : (int)(v - 0.5); }

// directories of a file path, created like mkpath (src/utils.h:86-116)
void mkpath(const std::string& file) {
    for (size_t p = file.find('/', 1); p != std::string::npos; p = file.find('/', p + 1)) {
        const std::string d = file.substr(0, p);
        if (mkdir(d.c_str(), 0775) != 0 && errno != EEXIST) fail("cannot create directory " + d);
    }
}

// stripbare (src/utils.h:14-32): file name without directory and extension
std::string stripbare(const std::string& path) {
    std::string s = path.substr(path.find_last_of('/') == std::string::npos ? 0 : path.find_last_of('/') + 1);
    const size_t dot = s.find_last_of('.');
    return dot == std::string::npos ? s : s.substr(0, dot);
}

// the set of names of a what= attribute (name_set, src/utils.h:39-84)
struct NameSet {
    std::set<std::string> names;
    bool all = false;
    explicit NameSet(const std::string& in) {
        if (in == "all") { all = true; return; }
        size_t b = 0;
        for (size_t i = 0; i <= in.size(); i++)
            if (i == in.size() || in[i] == ',') {
                if (i > b) names.insert(in.substr(b, i - b));
                b = i + 1;
            }
    }
    bool in(const std::string& w) const { return all || names.count(w) > 0; }
    bool explicitly(const std::string& w) const { return names.count(w) > 0; }
};

// ---------------------------------------------------------- bases ---
// Action::Init (src/Handlers/Action.cpp): Iterations and output=
struct Action : Handler {
    int init() override {
        if (const XmlAttr* a = node->attr("Iterations")) {
            start_iter = solver->iter;
            every_iter = solver->units.alt(a->value);
        } else {
            start_iter = solver->iter;
            every_iter = 0;
        }
        if (const XmlAttr* a = node->attr("output")) solver->set_output(a->value);
        return 0;
    }
    int type() const override { return ACTION; }
};

// Callback::Init (src/Handlers/Callback.cpp)
struct Callback : Handler {
    int init() override {
        start_iter = solver->iter;
        const XmlAttr* a = node->attr("Iterations");
        every_iter = a ? solver->units.alt(a->value) : 0;
        return 0;
    }
    int type() const override { return CALLBACK; }
};

// GenericAction (src/Handlers/GenericAction.cpp): runs the children; the
// callbacks with an interval are stacked on the solver until Unstack
struct GenericAction : Action {
    int stack = 0;
    int init() override {
        stack = 0;
        return Action::init();
    }
    int execute_internal() {
        stack = 0;
        for (auto& cp : node->children) {
            HandlerRef h = solver->make_handler(cp.get());
            if (!h) {
                solver->note("ERROR: Something wrong in " + node->name);
                return -1;
            }
            if (h->type() & CALLBACK) {
                if (h->every_iter != 0) {
                    solver->hands.push_back(h);
                    stack++;
                } else if (h->doit() != 0) {
                    solver->note("ERROR: Handler call error: " + cp->name);
                    return -1;
                }
            }
        }
        return 0;
    }
    void unstack() {
        while (stack > 0) {
            solver->hands.pop_back();
            stack--;
        }
    }
    int finish() override {
        if (stack > 0) unstack();
        return 0;
    }
};

// <CLBConfig> (src/Handlers/MainContainer.cpp): writes the configuration
// actually run, then the children
struct MainContainer : GenericAction {
    int init() override {
        GenericAction::init();
        XmlNode* run = solver->config->append_child("Run");
        run->set("model", solver->model->name);
        XmlNode* code = run->append_child("Code");
        code->set("version", "tclbx");
        code->set("precision", "double");
        code->set("cross", "GPU");
        return execute_internal();
    }
    int finish() override {
        unstack();
        return GenericAction::finish();
    }
};

// <Units> and other plain containers (src/Handlers/GenericContainer.cpp)
struct GenericContainer : GenericAction {
    int init() override {
        GenericAction::init();
        return execute_internal();
    }
    int finish() override {
        unstack();
        return GenericAction::finish();
    }
};

// <Container> (src/Handlers/acContainer.cpp): the children once; the
// callbacks they stacked come off when the container ends
struct AcContainer : GenericAction {
    int init() override {
        GenericAction::init();
        if (execute_internal()) return -1;
        unstack();
        return 0;
    }
};

// <Repeat Times> (src/Handlers/acRepeat.cpp): the children Times times.
// Each pass restarts the stack count, so only the last pass's interval
// callbacks come off at the end, as in the reference.
struct AcRepeat : GenericAction {
    int init() override {
        GenericAction::init();
        const XmlAttr* a = node->attr("Times");
        if (!a) {
            solver->note("ERROR: no Times parameter in Repeat");
            return -1;
        }
        const int times = std::atoi(a->value.c_str());
        for (int i = 0; i < times; i++)
            if (execute_internal()) return -1;
        return 0;
    }
};

// <Geometry> (src/Handlers/acGeometry.cpp)
struct AcGeometry : Action {
    int init() override {
        solver->load_geometry(*node);
        const GeomRegion& r = solver->geometry->region();
        if (solver->lattice) {
            Region o;
            o.dx = r.dx; o.dy = r.dy; o.dz = r.dz; o.nx = r.nx; o.ny = r.ny; o.nz = r.nz;
            solver->lattice->FlagOverwrite(solver->geometry->flags().data(), o);
            // Lattice::CutsOverwrite does nothing without cuts (src/Lattice.cu.Rt:1058)
            if (!solver->geometry->cuts().empty()) solver->lattice->CutsOverwrite(solver->geometry->cuts().data(), o);
        }
        return 0;
    }
};

// <Model> (src/Handlers/acModel.cpp): children, then Init
struct AcModel : GenericContainer {
    int init() override {
        int ret = GenericContainer::init();
        if (ret) return ret;
        if (solver->lattice) solver->lattice->Init();
        solver->iter = 0;
        return 0;
    }
};

// <Param name value [zone] [gauge]> (src/Handlers/acParam.cpp)
struct AcParam : Action {
    int init() override {
        Action::init();
        const std::string par = node->get("name"), zone = node->get("zone"), value = node->get("value"),
                          gauge = node->get("gauge");
        const XmlAttr* pa = context_attr("permissive");
        const bool permissive = pa && xml_as_bool(pa->value);
        int zone_number = -1;
        if (!zone.empty()) {
            auto it = solver->geometry->zones().find(zone);
            if (it == solver->geometry->zones().end()) {
                solver->note("ERROR: Unknown zone " + zone + " (found while setting parameter " + par + ")");
                return -1;
            }
            zone_number = it->second;
        }
        const double val = solver->units.alt(value);
        if (par.empty()) {
            if (!gauge.empty()) return 0;
            solver->note("ERROR: Setting name not specified in Param element");
            return -2;
        }
        const int sid = solver->model->setting_id(par);
        const int zid = solver->model->zonal_id(par);
        if (sid >= 0) {
            if (solver->lattice) solver->lattice->SetSetting(sid, val);
        } else if (zid >= 0) {
            if (solver->lattice) solver->lattice->SetZoneSetting(zid, zone_number, val);
        } else if (permissive) {
            solver->note((gauge.empty() ? "WARNING: Unknown setting " : "Unknown setting with gauge ") + par);
        } else {
            solver->note("ERROR: Unknown setting " + par);
            return -3;
        }
        return 0;
    }
};

// <Init> (src/Handlers/acInit.cpp)
struct AcInit : Action {
    int init() override {
        Action::init();
        if (solver->lattice) solver->lattice->Init();
        solver->iter = 0;
        return 0;
    }
};

// <Solve Iterations=..> (src/Handlers/acSolve.cpp): advance to the next
// iteration at which this element or a stacked callback is due, run the
// due callbacks, until this element is due or a callback stops it
struct AcSolve : GenericAction {
    int init() override {
        GenericAction::init();
        if (every_iter <= 0) {
            // the reference loops forever on a Solve without Iterations
            solver->note("ERROR: Solve needs a positive Iterations attribute");
            return -1;
        }
        if (execute_internal()) return -1;
        bool stop = false;
        do {
            const int my_next = next(solver->iter);
            int next_it = my_next;
            for (auto& h : solver->hands) {
                const int it = h->next(solver->iter);
                if (it > 0 && it < next_it) next_it = it;
            }
            const int steps = next_it;
            solver->iter += steps;
            int type = solver->iter_type;
            if (steps == my_next) type |= ITER_LASTGLOB;
            if (solver->lattice) solver->lattice->Iterate(steps, type);
            for (size_t i = 0; i < solver->hands.size(); i++) {
                auto h = solver->hands[i];
                if (h->now(solver->iter)) {
                    const int ret = h->doit();
                    if (ret == ITERATION_STOP) stop = true;
                    else if (ret != 0) return -1;
                }
            }
            if (stop) break;
        } while (!now(solver->iter));
        unstack();
        return 0;
    }
};

// <Log [name] Iterations> (src/Handlers/cbLog.cpp)
struct CbLog : Callback {
    std::string filename;
    int old_type = 0;
    int init() override {
        Callback::init();
        filename = solver->out_iter_file(node->get("name", "Log"), ".csv");
        if (solver->lattice) solver->init_log(filename);
        old_type = solver->iter_type;
        solver->iter_type |= ITER_LASTGLOB;
        return 0;
    }
    int doit() override {
        if (solver->lattice) solver->write_log(filename);
        return 0;
    }
    int finish() override {
        solver->iter_type = old_type;
        return 0;
    }
};

// <Stop <Global>Change|PercentChange|Above|Below=.. [Times]>
// (src/Handlers/cbStop.cpp)
struct CbStop : Callback {
    enum { CHANGE, PERCENT, ABOVE, BELOW };
    std::vector<int> what, kind;
    std::vector<double> limit, old;
    int times = 1, score = 0, old_type = 0;
    int init() override {
        Callback::init();
        const auto& G = solver->model->globals;
        for (int i = 0; i < (int)G.size(); i++) {
            const char* suffix[4] = {"Change", "PercentChange", "Above", "Below"};
            for (int k = 0; k < 4; k++)
                if (const XmlAttr* a = node->attr(G[i].name + suffix[k])) {
                    what.push_back(i);
                    kind.push_back(k);
                    limit.push_back(xml_as_double(a->value));
                    old.push_back(-12341234);
                }
        }
        if (what.empty()) {
            solver->note("ERROR: No *Change attribute in " + node->name);
            return -1;
        }
        if (const XmlAttr* a = node->attr("Times")) {
            times = xml_as_int(a->value);
            if (times < 1) {
                solver->note("ERROR: Minimal number for Times attribute is 1");
                return -1;
            }
        }
        score = 0;
        old_type = solver->iter_type;
        solver->iter_type |= ITER_LASTGLOB;
        return 0;
    }
    int doit() override {
        if (!solver->lattice) return 0;
        int any = 0;
        for (size_t i = 0; i < what.size(); i++) {
            const double v = solver->lattice->globals()[what[i]];
            switch (kind[i]) {
            case CHANGE: if (std::fabs(old[i] - v) > limit[i]) any++; break;
            case PERCENT: if (std::fabs((old[i] - v) / v) > limit[i]) any++; break;
            case ABOVE: if (v < limit[i]) any++; break;
            case BELOW: if (v > limit[i]) any++; break;
            }
            old[i] = v;
        }
        score = any ? 0 : score + 1;
        if (score >= times) {
            for (auto& o : old) o = -12341234;
            score = 0;
            solver->note("Stop at iteration " + std::to_string(solver->iter));
            return ITERATION_STOP;
        }
        return 0;
    }
    int finish() override {
        solver->iter_type = old_type;
        return 0;
    }
};

// <Failcheck [what] [dx..nz] Iterations> (src/Handlers/cbFailcheck.cpp).
// As in the reference: the region defaults to one node at the origin and
// the check runs once (the first time it is due); on a NaN the element's
// children run and the Solve stops.
struct CbFailcheck : Callback {
    GeomRegion reg;
    bool active = false;
    int init() override {
        Callback::init();
        active = false;
        reg = GeomRegion();
        if (const XmlAttr* a = node->attr("dx")) reg.dx = (int)solver->units.alt(a->value);
        if (const XmlAttr* a = node->attr("dy")) reg.dy = (int)solver->units.alt(a->value);
        if (const XmlAttr* a = node->attr("dz")) reg.dz = (int)solver->units.alt(a->value);
        if (const XmlAttr* a = node->attr("nx")) reg.nx = (int)solver->units.alt(a->value);
        if (const XmlAttr* a = node->attr("ny")) reg.ny = (int)solver->units.alt(a->value);
        if (const XmlAttr* a = node->attr("nz")) reg.nz = (int)solver->units.alt(a->value);
        return 0;
    }
    int doit() override {
        if (active || !solver->lattice) return 0;
        active = true;
        const NameSet comps(node->get("what", "all"));
        bool nan = false;
        const GeomRegion whole = GeomRegion{0, 0, 0, solver->nx, solver->ny, solver->nz};
        const GeomRegion r = whole.intersect(reg);
        for (int q = 0; q < (int)solver->model->quantities.size() && !nan; q++) {
            const QuantityDef& Q = solver->model->quantities[q];
            if (!comps.in(Q.name) || r.size() == 0) continue;
            std::vector<double> tmp((size_t)r.size() * (Q.vector ? 3 : 1));
            Region o;
            o.dx = r.dx; o.dy = r.dy; o.dz = r.dz; o.nx = r.nx; o.ny = r.ny; o.nz = r.nz;
            solver->lattice->GetQuantity(q, o, tmp.data(), 1.0);
            for (double v : tmp)
                if (std::isnan(v)) nan = true;
            if (nan) solver->note("Checking " + Q.name + " discovered NaN");
        }
        if (!nan) return 0;
        for (auto& cp : node->children) {
            HandlerRef h = solver->make_handler(cp.get());
            if (h) h->doit();
        }
        solver->note("Stopping due to NaN value");
        return ITERATION_STOP;
    }
};

// <VTK [name] [what] [dx..nz] Iterations> (src/Handlers/cbVTK.cpp)
struct CbVTK : Callback {
    std::string nm, what;
    GeomRegion reg;
    int init() override {
        Callback::init();
        nm = node->get("name", "VTK");
        what = node->get("what", "all");
        const GeomRegion total{0, 0, 0, solver->nx, solver->ny, solver->nz};
        reg = total;
        auto rd = [&](const char* a, int& v) {
            if (const XmlAttr* at = node->attr(a)) v = (int)solver->units.alt(at->value);
        };
        rd("dx", reg.dx);
        if (reg.dx < 0) { reg.dx = reg.nx + reg.dx; reg.nx = reg.nx - reg.dx; }
        rd("dy", reg.dy);
        if (reg.dy < 0) { reg.dy = reg.ny + reg.dy; reg.ny = reg.ny - reg.dy; }
        rd("dz", reg.dz);
        if (reg.dz < 0) { reg.dz = reg.nz + reg.dz; reg.nz = reg.nz - reg.dz; }
        rd("nx", reg.nx);
        if (reg.nx < 0) reg.nx = total.nx - reg.dx + reg.nx;
        rd("ny", reg.ny);
        if (reg.ny < 0) reg.ny = total.ny - reg.dy + reg.nz;  // (sic) as the reference
        rd("nz", reg.nz);
        if (reg.nz < 0) reg.nz = total.nz - reg.dz + reg.nz;
        reg = reg.intersect(total);
        if (reg.size() == 0) {
            solver->note("ERROR: VTK \"" + nm + "\" output has size 0");
            return -1;
        }
        return 0;
    }
    int doit() override { return solver->lattice ? solver->write_vtk(nm, what, reg) : 0; }
};

// <TXT [name] [what] [gzip] Iterations> (src/Handlers/cbTXT.cpp): every
// quantity as text, one lattice row of nx values per line
struct CbTXT : Callback {
    std::string nm, what;
    bool gzip = false;
    int init() override {
        Callback::init();
        nm = node->get("name", "TXT");
        what = node->get("what", "all");
        const XmlAttr* g = node->attr("gzip");
        gzip = g && xml_as_bool(g->value);
        return 0;
    }
    int doit() override { return solver->lattice ? solver->write_txt(nm, what, gzip) : 0; }
};

// <Average Iterations> (src/Handlers/cbAveraging.cpp): resetAverage when
// set up and every Iterations after
struct CbAverage : Callback {
    int init() override {
        Callback::init();
        if (solver->lattice) solver->lattice->resetAverage();
        return 0;
    }
    int doit() override {
        if (solver->lattice) solver->lattice->resetAverage();
        return 0;
    }
};

// <SaveBinary [file|filename] Iterations> (src/Handlers/cbSaveBinary.cpp):
// the current snapshot to <file>_0.pri (Lattice::saveSolution), in the
// reference's margin-block layout wherever a reference storage build holds
// the same stored values (FP64 and FP32 always), else the engine's layout
struct CbSaveBinary : Callback {
    std::string fn;
    int init() override {
        Callback::init();
        if (const XmlAttr* a = node->attr("file")) fn = solver->out_prefix() + "_" + a->value;
        else if (const XmlAttr* b = node->attr("filename")) fn = b->value;
        else fn = solver->out_iter_file("Save", "");
        return 0;
    }
    int doit() override {
        if (node->has("comp")) {
            solver->note("ERROR: SaveBinary comp= is not supported by this engine");
            return -1;
        }
        if (!solver->lattice) return 0;
        const std::string path = fn + "_0.pri";
        mkpath(path);
        return solver->lattice->Save(path, solver->lattice->case_layout()) == 0 ? 0 : -1;
    }
};

// <SaveCheckpoint keep=N|all> (src/Handlers/cbSaveCheckpoint.cpp): the state
// to <output>_checkpoint_<iter>_0.pri and a restart case
// <output>_restart_<iter>.xml -- the case file with a <LoadBinary> of that
// checkpoint before its first <Solve> (or the existing one re-pointed).
// Only the last `keep` checkpoints stay on disk (default 1; all: every one).
struct CbSaveCheckpoint : Callback {
    int keep = 1;
    std::deque<std::string> pri_, xml_;
    int init() override {
        Callback::init();
        if (const XmlAttr* a = node->attr("keep")) {
            if (a->value == "all") {
                keep = 0;
            } else {
                keep = xml_as_int(a->value);
                if (keep < 0) {
                    solver->note("ERROR: Keeping a negative no. of chckpnts not allowed, returning default behaviour.");
                    keep = 1;
                }
            }
        }
        return 0;
    }
    int doit() override {
        if (!solver->lattice) return 0;
        const std::string base = solver->out_collective_file("checkpoint", "");
        const std::string restart = solver->out_collective_file("restart", ".xml");
        const std::string pri = base + "_0.pri";
        if (solver->lattice->Save(pri, solver->lattice->case_layout()) != 0) return -1;
        write_restart(base, restart);
        if (keep != 0) {
            pri_.push_back(pri);
            xml_.push_back(restart);
            if ((int)pri_.size() > keep) {
                if (std::remove(pri_.front().c_str()) != 0)
                    solver->note("ERROR: Checkpoint file was not deleted: " + pri_.front());
                if (std::remove(xml_.front().c_str()) != 0)
                    solver->note("ERROR: Restart file was not deleted: " + xml_.front());
                pri_.pop_front();
                xml_.pop_front();
            }
        }
        return 0;
    }
    void write_restart(const std::string& base, const std::string& path) {
        std::unique_ptr<XmlNode> doc = solver->doc->clone();
        XmlNode* cfg = doc->child("CLBConfig");
        if (!cfg) return;
        if (XmlNode* lb = cfg->child("LoadBinary")) {
            lb->attrs.erase(std::remove_if(lb->attrs.begin(), lb->attrs.end(),
                                           [](const XmlAttr& a) { return a.name == "file"; }),
                            lb->attrs.end());
            lb->attrs.push_back({"file", base});
        } else {
            auto n = std::make_unique<XmlNode>();
            n->name = "LoadBinary";
            n->attrs.push_back({"file", base});
            n->parent = cfg;
            size_t at = 0;
            while (at < cfg->children.size() && cfg->children[at]->name != "Solve") at++;
            cfg->children.insert(cfg->children.begin() + (long)at, std::move(n));
        }
        std::ofstream f(path, std::ios::binary);
        f << xml_write(*doc);
    }
};

// <LoadBinary file|filename> (src/Handlers/acLoadBinary.cpp)
struct AcLoadBinary : Action {
    int init() override {
        Action::init();
        const XmlAttr* a = node->attr("file");
        if (!a) a = node->attr("filename");
        if (!a) {
            solver->note("ERROR: No file specified in LoadBinary");
            return -1;
        }
        if (node->has("comp")) {
            solver->note("ERROR: LoadBinary comp= is not supported by this engine");
            return -1;
        }
        if (solver->lattice && solver->lattice->Load(a->value + "_0.pri") != 0)
            fail("LoadBinary: cannot read " + a->value + "_0.pri");
        return 0;
    }
};

// elements that do nothing (src/Handlers/NullHandler.cpp)
struct NullHandler : Handler {
    int init() override { return 0; }
};

}  // namespace

// ----------------------------------------------------------- Handler ---
bool Handler::now(double it) const {
    if (every_iter == 0) return false;
    it -= start_iter;
    return std::floor(it / every_iter) > std::floor((it - 1) / every_iter);
}

int Handler::next(double it) const {
    if (every_iter == 0) return -1;
    it -= start_iter;
    const int k = (int)std::floor(it / every_iter);
    return (int)(-std::floor(-(k + 1) * every_iter) - it);
}

const XmlAttr* Handler::context_attr(const char* name) const {
    for (const XmlNode* n = node; n; n = n->parent)
        if (const XmlAttr* a = n->attr(name)) return a;
    return nullptr;
}

// ------------------------------------------------------------ Solver ---
Solver::Solver(const std::string& m, const std::string& st) {
    model = model_by_name(m);
    if (!model) fail("unknown model " + m);
    storage = storage_by_name(st);
    if (storage < 0) fail("unknown storage format " + st);
    t0_ = walltime();
}

Solver::~Solver() = default;

double Solver::walltime() const {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count() - t0_;
}

void Solver::load_file(const std::string& path) {
    doc = xml_load_file(path);
    conffile = path;
    prepare();
}

void Solver::load_string(const std::string& text, const std::string& name) {
    doc = xml_parse(text);
    conffile = name;
    prepare();
}

// main.cpp:246-375: deprecated <Params>, Units, the lattice size
void Solver::prepare() {
    config = doc->child("CLBConfig");
    if (!config) fail(conffile + ": No \"CLBConfig\" element");
    set_output("");
    // <Params a="1" b-zone="2"/> becomes <Param name="a" value="1"/>... and
    // makes the case permissive (main.cpp:250-279)
    std::vector<XmlNode*> params;
    std::vector<XmlNode*> stack{config};
    while (!stack.empty()) {
        XmlNode* n = stack.back();
        stack.pop_back();
        for (auto& c : n->children) {
            if (c->name == "Params") params.push_back(c.get());
            stack.push_back(c.get());
        }
    }
    if (!params.empty()) {
        config->set("permissive", "true");
        note("WARNING: deprecated Params elements changed to Param");
        for (XmlNode* node : params) {
            XmlNode* par = node->parent;
            size_t pos = 0;
            while (par->children[pos].get() != node) pos++;
            const std::string gauge = node->get("gauge");
            std::vector<std::unique_ptr<XmlNode>> made;
            for (auto& a : node->attrs) {
                if (a.name == "gauge") continue;
                auto p = std::make_unique<XmlNode>();
                p->name = "Param";
                p->parent = par;
                std::string name = a.name, zone;
                const size_t dash = name.find('-');
                if (dash != std::string::npos) { zone = name.substr(dash + 1); name = name.substr(0, dash); }
                p->set("name", name);
                p->set("value", a.value);
                if (!zone.empty()) p->set("zone", zone);
                if (!gauge.empty()) p->set("gauge", gauge);
                made.push_back(std::move(p));
            }
            par->children.erase(par->children.begin() + pos);
            for (size_t k = 0; k < made.size(); k++) par->children.insert(par->children.begin() + pos + k, std::move(made[k]));
        }
    }
    // readUnits (main.cpp:31-62)
    if (XmlNode* u = config->child("Units")) {
        int i = 1;
        for (auto& c : u->children) {
            if (c->name != "Param") continue;
            if (!c->has("value")) fail("Value not provided in Param in Units");
            if (!c->has("gauge")) fail("Gauge not provided in Param in Units");
            units.set_gauge(c->get("name", "unnamed" + std::to_string(i)), c->get("value"), c->get("gauge"));
            i++;
        }
        units.make_gauge();
    }
    XmlNode* geom = config->child("Geometry");
    if (!geom) fail(conffile + ": No \"Geometry\" element");
    nx = myround(units.alt(geom->get("nx"), 1));
    ny = myround(units.alt(geom->get("ny"), 1));
    nz = myround(units.alt(geom->get("nz"), 1));
    if (nx <= 0 || ny <= 0 || nz <= 0) fail("lattice size must be positive");
    if (model->dim == 2 && nz != 1) fail(model->name + " is a 2D model: nz must be 1");
    geometry = std::make_unique<Geometry>(*model, nx, ny, nz, units);
}

void Solver::rasterize() {
    if (!config) fail("no case loaded");
    for (auto& c : config->children)
        if (c->name == "Geometry") load_geometry(*c);
}

void Solver::load_geometry(XmlNode& g) {
    const size_t k = geometry->notes().size();
    geometry->load(g);
    for (size_t i = k; i < geometry->notes().size(); i++) note(geometry->notes()[i]);
}

void Solver::run() {
    if (!config) fail("no case loaded");
    // a fresh NodeType map: the Geometry elements draw it as they run
    geometry = std::make_unique<Geometry>(*model, nx, ny, nz, units);
    lattice = std::make_unique<Lattice>(model, nx, ny, nz, storage);
    iter = 0;
    iter_type = ITER_NORM;
    HandlerRef main = make_handler(config);
    if (!main) {
        std::string last = messages.empty() ? "" : messages.back();
        hands.clear();
        fail("case failed: " + last);
    }
    main.reset();
    hands.clear();
}

void Solver::check() {
    if (!config) fail("no case loaded");
    geometry = std::make_unique<Geometry>(*model, nx, ny, nz, units);
    lattice.reset();
    iter = 0;
    iter_type = ITER_NORM;
    HandlerRef main = make_handler(config);
    if (!main) {
        std::string last = messages.empty() ? "" : messages.back();
        hands.clear();
        fail("case failed: " + last);
    }
    main.reset();
    hands.clear();
}

HandlerRef Solver::make_handler(XmlNode* n) {
    Handler* h = nullptr;
    const std::string& e = n->name;
    if (e == "CLBConfig") h = new MainContainer;
    else if (e == "Units") h = new GenericContainer;
    else if (e == "Geometry") h = new AcGeometry;
    else if (e == "Model") h = new AcModel;
    else if (e == "Param") h = new AcParam;
    else if (e == "Init") h = new AcInit;
    else if (e == "Solve") h = new AcSolve;
    else if (e == "Log") h = new CbLog;
    else if (e == "Stop") h = new CbStop;
    else if (e == "Failcheck") h = new CbFailcheck;
    else if (e == "VTK") h = new CbVTK;
    else if (e == "TXT") h = new CbTXT;
    else if (e == "Container") h = new AcContainer;
    else if (e == "Repeat") h = new AcRepeat;
    else if (e == "SaveBinary") h = new CbSaveBinary;
    else if (e == "SaveCheckpoint") h = new CbSaveCheckpoint;
    else if (e == "Average") h = new CbAverage;
    else if (e == "LoadBinary") h = new AcLoadBinary;
    else if (e == "Run") h = new NullHandler;
    else {
        note("ERROR: Unknown Handler: " + e);
        h = new NullHandler;
    }
    h->node = n;
    h->solver = this;
    int ret;
    try {
        ret = h->init();
    } catch (...) {
        delete h;
        throw;
    }
    if (ret) {
        delete h;
        return nullptr;
    }
    // the last reference runs finish() (src/Handlers.h:35-50)
    return HandlerRef(h, [](Handler* p) {
        p->finish();
        delete p;
    });
}

void Solver::set_output(const std::string& prefix) { outpath = prefix + stripbare(conffile); }

std::string Solver::out_iter_file(const std::string& name, const std::string& suffix) const {
    char num[32];
    std::snprintf(num, sizeof num, "_P%02d_%08d", 0, iter);
    std::string p = root_ + outpath + "_" + name + num + suffix;
    mkpath(p);
    return p;
}

std::string Solver::out_collective_file(const std::string& name, const std::string& suffix) const {
    char num[32];
    std::snprintf(num, sizeof num, "_%08d", iter);
    std::string p = root_ + outpath + "_" + name + num + suffix;
    mkpath(p);
    return p;
}

std::string Solver::out_global_file(const std::string& name, const std::string& suffix) const {
    std::string p = root_ + outpath + "_" + name + "_P00" + suffix;
    mkpath(p);
    return p;
}

std::vector<int> Solver::log_settings() const {
    std::vector<int> r;
    for (int i = 0; i < (int)model->settings.size(); i++)
        if (!model->settings[i].extension) r.push_back(i);
    return r;
}

// initLog (src/Solver.cpp.Rt:126-172)
void Solver::init_log(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "wt");
    if (!f) fail("cannot open " + path);
    std::fprintf(f, "\"Iteration\",\"Time_si\",\"Walltime\",\"Optimization\"");
    for (int i : log_settings()) std::fprintf(f, ",\"%s\",\"%s_si\"", model->settings[i].name.c_str(), model->settings[i].name.c_str());
    for (auto& z : model->zonal)
        for (auto& zone : geometry->zones())
            std::fprintf(f, ",\"%s-%s\",\"%s-%s_si\"", z.name.c_str(), zone.first.c_str(), z.name.c_str(),
                         zone.first.c_str());
    for (auto& g : model->globals) std::fprintf(f, ",\"%s\",\"%s_si\"", g.name.c_str(), g.name.c_str());
    for (const char* s : {"dx", "dt", "dm"}) std::fprintf(f, ",\"%s_si\"", s);
    std::fprintf(f, "\n");
    std::fclose(f);
}

// writeLog (src/Solver.cpp.Rt:178-206): values and SI values (value / the
// lattice size of its unit)
void Solver::write_log(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "at");
    if (!f) fail("cannot open " + path);
    std::fprintf(f, "%d, %.13le, %.13le, %d", iter, 1 / units.alt("s") * iter, walltime(), 0);
    for (int i : log_settings()) {
        const double v = lattice->GetSetting(i);
        std::fprintf(f, ", %.13le, %.13le", v, v * (1 / units.alt(model->settings[i].unit)));
    }
    for (size_t i = 0; i < model->zonal.size(); i++) {
        const double sc = 1 / units.alt(model->zonal[i].unit);
        for (auto& zone : geometry->zones()) {
            const double v = lattice->GetZoneSetting((int)i, zone.second);
            std::fprintf(f, ", %.13le, %.13le", v, v * sc);
        }
    }
    const auto& g = lattice->globals();
    for (size_t i = 0; i < model->globals.size(); i++)
        std::fprintf(f, ", %.13le, %.13le", g[i], g[i] * (1 / units.alt(model->globals[i].unit)));
    for (const char* s : {"m", "s", "kg"}) std::fprintf(f, ", %.13le", 1 / units.alt(s));
    std::fprintf(f, "\n");
    std::fclose(f);
}

namespace {
// base64 of src/vtkOutput.cpp:6-36: 3 bytes -> 4 characters, one line
void b64(FILE* f, const void* data, size_t len) {
    static const char* C = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    const unsigned char* p = (const unsigned char*)data;
    std::string out;
    out.reserve((len + 2) / 3 * 4);
    for (size_t i = 0; i < len; i += 3) {
        const size_t n = len - i < 3 ? len - i : 3;
        unsigned w = 0;
        for (size_t k = 0; k < 3; k++) w = (w << 8) + (k < n ? p[i + k] : 0);
        char c[4];
        for (int k = 0; k < 4; k++) { c[3 - k] = C[w & 0x3F]; w >>= 6; }
        for (size_t k = n; k < 3; k++) c[k + 1] = '=';
        out.append(c, 4);
    }
    std::fwrite(out.data(), 1, out.size(), f);
}

void vtk_field(FILE* f, FILE* fp, const char* name, const void* data, size_t bytes, const char* tp, int comp) {
    std::fprintf(f, "<DataArray type=\"%s\" Name=\"%s\" format=\"binary\" encoding=\"base64\" NumberOfComponents=\"%d\">\n",
                 tp, name, comp);
    const int len = (int)bytes;
    b64(f, &len, sizeof(int));
    b64(f, data, bytes);
    std::fprintf(f, "\n</DataArray>\n");
    if (fp)
        std::fprintf(fp, "<PDataArray type=\"%s\" Name=\"%s\" format=\"binary\" encoding=\"base64\" NumberOfComponents=\"%d\"/>\n",
                     tp, name, comp);
}
}  // namespace

// vtkWriteLattice (src/vtkLattice.cpp:9-57) with the file layout of
// vtkFileOut (src/vtkOutput.cpp): <name>.vti with the selected NodeType
// groups (UInt8) and quantities (Float64, in SI units), and the .pvti that
// lists it
int Solver::write_vtk(const std::string& name, const std::string& what_s, const GeomRegion& reg) {
    const NameSet what(what_s);
    const std::string fn = out_iter_file(name, ".vti");
    std::string pfn = fn.substr(0, fn.size() - 4) + ".pvti";
    FILE* f = std::fopen(fn.c_str(), "w");
    if (!f) fail("Could not open vtk file " + fn);
    FILE* fp = std::fopen(pfn.c_str(), "w");
    if (!fp) { std::fclose(f); fail("Could not open (p)vtk file " + pfn); }
    const double spacing = 1 / units.alt("m");
    const char* sel = "Scalars=\"rho\" Vectors=\"velocity\"";
    char ext[160];
    std::snprintf(ext, sizeof ext, "%d %d %d %d %d %d", reg.dx, reg.dx + reg.nx, reg.dy, reg.dy + reg.ny, reg.dz,
                  reg.dz + reg.nz);
    std::fprintf(f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
    std::fprintf(f, "<ImageData WholeExtent=\"%s\" Origin=\"%lg %lg %lg\" Spacing=\"%lg %lg %lg\">\n", ext, 0.0, 0.0,
                 0.0, spacing, spacing, spacing);
    std::fprintf(f, "<Piece Extent=\"%s\">\n", ext);
    std::fprintf(f, "<CellData %s>\n", sel);
    std::fprintf(fp, "<?xml version=\"1.0\"?>\n<VTKFile type=\"PImageData\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
    std::fprintf(fp, "<PImageData WholeExtent=\"%s\" Origin=\"%lg %lg %lg\" Spacing=\"%lg %lg %lg\">\n", ext, 0.0, 0.0,
                 0.0, spacing, spacing, spacing);
    const std::string base = fn.substr(fn.find_last_of('/') == std::string::npos ? 0 : fn.find_last_of('/') + 1);
    std::fprintf(fp, "<Piece Extent=\"%s\" Source=\"%s\"/>\n", ext, base.c_str());
    std::fprintf(fp, "<PCellData %s>\n", sel);

    const size_t n = (size_t)reg.size();
    Region o;
    o.dx = reg.dx; o.dy = reg.dy; o.dz = reg.dz; o.nx = reg.nx; o.ny = reg.ny; o.nz = reg.nz;
    std::vector<uint16_t> flags(n);
    lattice->GetFlags(o, flags.data());
    if (what.explicitly("flag")) vtk_field(f, fp, "flag", flags.data(), n * 2, "UInt16", 1);
    std::vector<unsigned char> small(n);
    for (auto& g : model->groups) {
        const bool save = g.name != "NONE" && g.name != "ALL";
        if ((what.all && save) || what.explicitly(g.name)) {
            for (size_t i = 0; i < n; i++) small[i] = (unsigned char)((flags[i] & g.mask) >> g.shift);
            vtk_field(f, fp, g.name.c_str(), small.data(), n, "UInt8", 1);
        }
    }
    for (int q = 0; q < (int)model->quantities.size(); q++) {
        const QuantityDef& Q = model->quantities[q];
        if (!what.in(Q.name)) continue;
        const int comp = Q.vector ? 3 : 1;
        std::vector<double> tmp(n * comp);
        lattice->GetQuantity(q, o, tmp.data(), 1 / units.alt(Q.unit));
        vtk_field(f, fp, Q.name.c_str(), tmp.data(), n * comp * sizeof(double), "Float64", comp);
    }
    std::fprintf(f, "</CellData>\n</Piece>\n</ImageData>\n</VTKFile>\n");
    std::fprintf(fp, "</PCellData>\n</PImageData>\n</VTKFile>\n");
    std::fclose(f);
    std::fclose(fp);
    return 0;
}

// The text output of src/vtkLattice.cpp:85-160 (txtWriteLattice): an _info
// file with the unit scales and the size, then one <prefix>_<quantity>.txt
// per selected quantity in the output unit, %.16lg values (a vector as its
// three components), a newline after every nx values.  gzip writes each as
// <file>.gz (zlib in process, where the reference pipes through gzip).
int Solver::write_txt(const std::string& name, const std::string& what_s, bool gzip) {
    const NameSet what(what_s);
    const std::string base = out_iter_file(name, "");
    const std::string info = base + "_info.txt";
    FILE* f = std::fopen(info.c_str(), "w");
    if (!f) {
        note("ERROR: Cannot open file: " + info);
        return -1;
    }
    const size_t n = (size_t)nx * ny * nz;
    std::fprintf(f, "dx: %lg\n", 1 / units.alt("m"));
    std::fprintf(f, "dt: %lg\n", 1 / units.alt("s"));
    std::fprintf(f, "dm: %lg\n", 1 / units.alt("kg"));
    std::fprintf(f, "dT: %lg\n", 1 / units.alt("K"));
    std::fprintf(f, "size: %d\n", (int)n);
    std::fprintf(f, "NX: %d\nNY: %d\nNZ: %d\n", nx, ny, nz);
    std::fclose(f);
    Region o;
    o.dx = 0; o.dy = 0; o.dz = 0; o.nx = nx; o.ny = ny; o.nz = nz;
    for (int q = 0; q < (int)model->quantities.size(); q++) {
        const QuantityDef& Q = model->quantities[q];
        if (!what.in(Q.name)) continue;
        const std::string fn = base + "_" + Q.name + ".txt";
        const int comp = Q.vector ? 3 : 1;
        std::vector<double> tmp(n * comp);
        lattice->GetQuantity(q, o, tmp.data(), 1 / units.alt(Q.unit));
        std::string text;
        char num[40];
        for (size_t i = 0; i < n; i++) {
            for (int c = 0; c < comp; c++) {
                if (c) text += ' ';
                text.append(num, (size_t)std::snprintf(num, sizeof num, "%.16lg", tmp[i * comp + c]));
            }
            text += (i + 1) % (size_t)nx == 0 ? '\n' : ' ';
        }
        bool ok;
        if (gzip) {
            gzFile z = gzopen((fn + ".gz").c_str(), "wb");
            ok = z != nullptr && gzwrite(z, text.data(), (unsigned)text.size()) == (int)text.size();
            if (z && gzclose(z) != Z_OK) ok = false;
        } else {
            FILE* out = std::fopen(fn.c_str(), "w");
            ok = out != nullptr && std::fwrite(text.data(), 1, text.size(), out) == text.size();
            if (out && std::fclose(out) != 0) ok = false;
        }
        if (!ok) {
            note("ERROR: Cannot write file: " + fn + (gzip ? ".gz" : ""));
            return -1;
        }
    }
    return 0;
}

}  // namespace tclbx
