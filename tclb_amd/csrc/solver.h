// solver.h -- runs a TCLB case file on the engine.
//
// TCLB's executable reads the case file, solves the unit gauges, sizes the
// lattice from <Geometry nx ny nz> and then executes the elements of
// <CLBConfig> as handlers (src/main.cpp:246-418, src/Handlers/*): actions
// (Geometry, Model, Param, Init, Solve, LoadBinary, ...) run once when
// reached, callbacks (Log, Stop, Failcheck, VTK, SaveBinary) run every
// `Iterations` steps of the enclosing Solve, which advances the lattice to
// the next iteration at which any of them is due (src/Handlers/acSolve.cpp).
// This class does the same over a tclbx::Lattice.  The outputs keep the
// reference's names (<output><case>_<Name>_P00_<iter>.<ext>) and formats:
// the Log CSV of src/Solver.cpp.Rt:120-206 and the VTI image data of
// src/vtkLattice.cpp / src/vtkOutput.cpp.
#pragma once
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "geometry.h"
#include "lattice.h"
#include "model.h"
#include "units.h"
#include "xml.h"

namespace tclbx {

class Solver;

// one element of the case file (src/Handlers/vHandler.h:25-105)
struct Handler {
    enum { CALLBACK = 0x01, ACTION = 0x02, DESIGN = 0x04 };
    XmlNode* node = nullptr;
    Solver* solver = nullptr;
    int start_iter = 0;
    double every_iter = 0;
    virtual ~Handler() = default;
    virtual int init() = 0;           // 0 ok, nonzero error
    virtual int doit() { return 0; }  // 0 ok, ITERATION_STOP, or error
    virtual int finish() { return 0; }
    virtual int type() const { return 0; }
    // due at iteration `iter` / steps until the next due iteration
    bool now(double iter) const;
    int next(double iter) const;
    // attribute of this element or the nearest ancestor that has it
    const XmlAttr* context_attr(const char* name) const;
};
using HandlerRef = std::shared_ptr<Handler>;

constexpr int ITERATION_STOP = 1;  // src/CommonHandler.h:10

class Solver {
public:
    Solver(const std::string& model, const std::string& storage);
    ~Solver();
    // reads a case file (the path also names the outputs) / case text
    void load_file(const std::string& path);
    void load_string(const std::string& text, const std::string& conffile);
    // directory prepended to every output path (default: none)
    void set_output_root(const std::string& dir) { root_ = dir; }
    // host only: the Units and the NodeType map of every <Geometry> element
    void rasterize();
    // the whole case: lattice on the GPU, every handler in order
    void run();
    // the same walk without a lattice (no GPU, no output files): every
    // handler is made and initialised, its attributes, units, settings and
    // zones resolved and the geometry drawn, and the Solve loops step
    // through their iterations and callbacks without computing anything.
    // Throws like run(); the handlers' errors are in `messages`.
    void check();

    // --- state the handlers use
    const ModelDef* model;
    int storage;
    Units units;
    std::unique_ptr<XmlNode> doc;
    XmlNode* config = nullptr;
    std::unique_ptr<Geometry> geometry;
    std::unique_ptr<Lattice> lattice;
    int nx = 1, ny = 1, nz = 1;
    int iter = 0;
    int iter_type = ITER_NORM;
    std::vector<HandlerRef> hands;
    std::string conffile, outpath;
    std::vector<std::string> messages;  // what the reference would print as notices/errors

    HandlerRef make_handler(XmlNode* n);
    void set_output(const std::string& prefix);
    std::string out_iter_file(const std::string& name, const std::string& suffix) const;
    std::string out_global_file(const std::string& name, const std::string& suffix) const;
    // <output><case>_<name>_<iter><suffix> (outIterCollectiveFile, src/Solver.h.Rt:126-129)
    std::string out_collective_file(const std::string& name, const std::string& suffix) const;
    // the output root followed by the output path (<output><case>)
    std::string out_prefix() const { return root_ + outpath; }
    void init_log(const std::string& path);
    void write_log(const std::string& path);
    int write_vtk(const std::string& name, const std::string& what, const GeomRegion& reg);
    int write_txt(const std::string& name, const std::string& what, bool gzip);
    double walltime() const;
    void note(const std::string& m) { messages.push_back(m); }
    // draws a <Geometry> element; its notices join `messages`
    void load_geometry(XmlNode& g);
    // log columns: settings of the model that the Log lists
    std::vector<int> log_settings() const;

private:
    void prepare();
    std::string root_;
    double t0_ = 0;
};

}  // namespace tclbx
