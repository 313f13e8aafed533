// units.h -- TCLB's unit system for case files (src/unit.h, src/unit.cpp).
//
// A quantity is a value times a product of powers of nine base units
// m, s, kg, K, x, y, z, A, t.  A case file's <Units><Param value=".."
// gauge=".."/></Units> entries state "value [with units] = gauge [lattice
// units]"; solving the log-linear system of those equations gives the
// lattice size of every base unit, and alt("...") converts any text such as
// "0.02m2/s", "10cm+2" or "1e-6" into lattice units.  The parsing rules,
// the named units and prefixes and the elimination used to solve the
// gauges are those of the reference, so the same text converts to the same
// double.
#pragma once
#include <map>
#include <string>

namespace tclbx {

constexpr int N_BASE_UNITS = 9;

struct UnitVal {
    double val = 0;
    double uni[N_BASE_UNITS] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    UnitVal() = default;
    UnitVal(double v) : val(v) {}  // NOLINT: numbers convert implicitly, as in the reference
    static UnitVal base(int k);
    UnitVal operator*(const UnitVal& o) const;
    UnitVal operator/(const UnitVal& o) const;
    UnitVal pow(double p) const;
};

class Units {
public:
    Units();
    // text with units -> value in SI (unit scales 1) / in lattice units
    UnitVal read_text(const std::string& s) const;
    double alt(const UnitVal& v) const;
    // sums the +/- separated terms of `s`, each converted to lattice units
    double alt(const std::string& s) const;
    double alt(const std::string& s, double def) const { return s.empty() ? def : alt(s); }
    // one <Units><Param> equation: value / gauge = 1 lattice unit
    void set_gauge(const std::string& name, const std::string& value, const std::string& gauge);
    // solves the gauges for the scales of the base units
    void make_gauge();
    double scale(int k) const { return scale_[k]; }

private:
    UnitVal read_unit_one(const std::string& s) const;
    UnitVal read_unit_alpha(const std::string& s, double p) const;
    UnitVal read_unit(const std::string& s) const;
    std::map<std::string, UnitVal> units_;
    std::map<std::string, UnitVal> gauge_;
    double scale_[N_BASE_UNITS];
};

}  // namespace tclbx
