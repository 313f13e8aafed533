// units.cpp -- see units.h.  Behaviour follows src/unit.cpp and src/unit.h.
#include "units.h"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <stdexcept>
#include <vector>

namespace tclbx {

namespace {
const char* const kBase[N_BASE_UNITS] = {"m", "s", "kg", "K", "x", "y", "z", "A", "t"};
}

UnitVal UnitVal::base(int k) {
    UnitVal u(1.0);
    u.uni[k] = 1;
    return u;
}

UnitVal UnitVal::operator*(const UnitVal& o) const {
    UnitVal r(val * o.val);
    for (int i = 0; i < N_BASE_UNITS; i++) r.uni[i] = uni[i] + o.uni[i];
    return r;
}

UnitVal UnitVal::operator/(const UnitVal& o) const {
    UnitVal r(val / o.val);
    for (int i = 0; i < N_BASE_UNITS; i++) r.uni[i] = uni[i] - o.uni[i];
    return r;
}

UnitVal UnitVal::pow(double p) const {
    UnitVal r(std::pow(val, p));
    for (int i = 0; i < N_BASE_UNITS; i++) r.uni[i] = uni[i] * p;
    return r;
}

// base units, derived units and prefixed units of src/unit.cpp:68-103
Units::Units() {
    for (int i = 0; i < N_BASE_UNITS; i++) {
        scale_[i] = 1;
        units_[kBase[i]] = UnitVal::base(i);
    }
    units_["N"] = read_text("1kgm/s2");
    units_["Pa"] = read_text("1N/m2");
    units_["J"] = read_text("1Nm");
    units_["W"] = read_text("1J/s");
    units_["V"] = read_text("1kgm2/t3/A");
    units_["C"] = read_text("1tA");
    units_["nm"] = read_text("1e-9m");
    units_["um"] = read_text("1e-6m");
    units_["mm"] = read_text("1e-3m");
    units_["cm"] = read_text("1e-2m");
    units_["km"] = read_text("1e+3m");
    units_["h"] = read_text("3600s");
    units_["ns"] = read_text("1e-9s");
    units_["us"] = read_text("1e-6s");
    units_["ms"] = read_text("1e-3s");
    units_["g"] = read_text("1e-3kg");
    units_["mg"] = read_text("1e-6kg");
    units_["d"] = UnitVal(std::atan(1.0) * 4.0 / 180.);
    units_["%"] = UnitVal(1. / 100.);
    units_["An"] = UnitVal(6.022 * 100000000000000000000000.);
}

UnitVal Units::read_unit_one(const std::string& s) const {
    auto it = units_.find(s);
    return it == units_.end() ? UnitVal(0.0) : it->second;
}

// a run of letters: either its first letter is a unit followed by units, or
// its first two letters are; when both readings exist, a leading "m" is milli
UnitVal Units::read_unit_alpha(const std::string& s, double p) const {
    UnitVal r1 = read_unit_one(s.substr(0, 1));
    if (s.size() < 2) return r1.pow(p);
    r1 = r1 * read_unit_alpha(s.substr(1), p);
    UnitVal r2 = read_unit_one(s.substr(0, 2));
    if (s.size() > 2) r2 = r2 * read_unit_alpha(s.substr(2), p);
    else r2 = r2.pow(p);
    if (r1.val == 0) return r2.val == 0 ? UnitVal(0.0) : r2;
    if (r2.val == 0) return r1;
    if (s[0] == 'm') return r2;
    throw std::runtime_error("ambiguous unit: \"" + s + "\"");
}

// letters with an optional power ("m2", "s"), multiplied; after a "/" every
// later factor divides
UnitVal Units::read_unit(const std::string& s) const {
    UnitVal ret(1.0);
    int sign = 1;
    size_t i = 0;
    auto at = [&](size_t k) { return k < s.size() ? s[k] : '\0'; };
    while (at(i)) {
        const size_t start = i;
        size_t j = i;
        while (std::isalpha((unsigned char)at(i))) i++;
        size_t k = i;
        while (std::isdigit((unsigned char)at(i)) || at(i) == '.') i++;
        size_t l = i;
        double p = 1;
        if (l > k) p = std::atof(s.substr(k, l - k).c_str());
        UnitVal last = k > j ? read_unit_alpha(s.substr(j, k - j), p) : UnitVal(1.0);
        ret = sign > 0 ? ret * last : ret / last;
        j = i;
        while (!std::isalnum((unsigned char)at(i)) && at(i) > 0) i++;
        if (i - j > 1) throw std::runtime_error("too many non-alphanumeric characters in units: \"" + s.substr(j, i - j) + "\"");
        if (i - j == 1) {
            if (s[j] == '/') sign = -1;
            else throw std::runtime_error(std::string("only \"/\" allowed in units: \"") + s[j] + "\"");
        }
        // a byte that is neither a letter, a digit nor ASCII punctuation
        // makes no progress (the reference would loop forever on it)
        if (i == start) throw std::runtime_error("unreadable character in units: \"" + s + "\"");
    }
    return ret;
}

// a leading number ([-+0-9.eE]*) times the units that follow it
UnitVal Units::read_text(const std::string& s) const {
    size_t i = 0;
    while (i < s.size()) {
        const char c = s[i];
        if (c == '-' || c == '+' || (c >= '0' && c <= '9') || c == '.' || c == 'e' || c == 'E') i++;
        else break;
    }
    UnitVal ret = read_unit(s.substr(i));
    if (i > 0) ret = ret * UnitVal(std::atof(s.substr(0, i).c_str()));
    return ret;
}

double Units::alt(const UnitVal& v) const {
    double r = v.val;
    for (int i = 0; i < N_BASE_UNITS; i++) r *= std::pow(scale_[i], v.uni[i]);
    return r;
}

// terms split before every + or - that is not an exponent sign
double Units::alt(const std::string& s) const {
    double ret = 0;
    size_t i = 0, j = 0;
    auto at = [&](size_t k) { return k < s.size() ? s[k] : '\0'; };
    while (at(i)) {
        const char c = at(j);
        if (c == '-' || c == '+' || c == '\0') {
            if (j > i) ret += alt(read_text(s.substr(i, j - i)));
            i = j;
        } else if ((c == 'e' || c == 'E') && (at(j + 1) == '-' || at(j + 1) == '+')) {
            j++;
        }
        j++;
    }
    return ret;
}

void Units::set_gauge(const std::string& name, const std::string& value, const std::string& gauge) {
    gauge_[name] = read_text(value) / read_text(gauge) / UnitVal(1.0);
}

// log(scale) solves one equation per gauge (sum_j uni_j log s_j = -log val);
// base units no gauge mentions get the equation log s_j = 0.  Elimination
// with a row swap only on a zero pivot, as src/unit.cpp:13-49.
void Units::make_gauge() {
    const int n = N_BASE_UNITS;
    std::vector<double> A(n * n, 0.0), b(n, 0.0), x(n, 0.0);
    auto IND = [](int r, int c) { return c * N_BASE_UNITS + r; };
    int rows = 0;
    for (auto& g : gauge_) {
        if (rows >= n) throw std::runtime_error("gauge variables over-constructed");
        for (int j = 0; j < n; j++) A[IND(rows, j)] = g.second.uni[j];
        b[rows] = std::log(g.second.val);
        rows++;
    }
    for (int j = 0; j < n; j++) {
        bool used = false;
        for (int l = 0; l < rows; l++) if (A[IND(l, j)] != 0) used = true;
        if (!used) {
            if (rows >= n) throw std::runtime_error("gauge variables over-constructed");
            A[IND(rows, j)] = 1;
            b[rows] = 0;
            rows++;
        }
    }
    if (rows < n) throw std::runtime_error("gauge variables under-constructed");
    auto is_zero = [](double v) { return std::fabs(v) < 1e-10; };
    for (int k = 0; k < n - 1; k++) {
        if (is_zero(A[IND(k, k)]))
            for (int i = k + 1; i < n; i++)
                if (!is_zero(A[IND(i, k)])) {
                    for (int j = 0; j < n; j++) std::swap(A[IND(k, j)], A[IND(i, j)]);
                    std::swap(b[k], b[i]);
                    break;
                }
        for (int i = k + 1; i < n; i++) {
            const double m = A[IND(i, k)] / A[IND(k, k)];
            for (int j = 0; j < n; j++) A[IND(i, j)] -= m * A[IND(k, j)];
            b[i] -= m * b[k];
        }
    }
    for (int i = n - 1; i >= 0; i--) {
        double sum = 0.0;
        for (int j = n - 1; j >= i + 1; j--) sum += A[IND(i, j)] * x[j];
        x[i] = (b[i] - sum) / A[IND(i, i)];
    }
    for (int j = 0; j < n; j++) scale_[j] = std::exp(-x[j]);
}

}  // namespace tclbx
