// xml.cpp -- see xml.h.
#include "xml.h"

#include <climits>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace tclbx {

const XmlAttr* XmlNode::attr(const std::string& n) const {
    for (auto& a : attrs) if (a.name == n) return &a;
    return nullptr;
}

std::string XmlNode::get(const std::string& n, const std::string& def) const {
    const XmlAttr* a = attr(n);
    return a ? a->value : def;
}

XmlNode* XmlNode::child(const std::string& n) const {
    for (auto& c : children) if (c->name == n) return c.get();
    return nullptr;
}

XmlNode* XmlNode::find_child_by_attribute(const std::string& elem, const std::string& a, const std::string& v) const {
    for (auto& c : children) {
        if (c->name != elem) continue;
        const XmlAttr* at = c->attr(a);
        if (at && at->value == v) return c.get();
    }
    return nullptr;
}

std::unique_ptr<XmlNode> XmlNode::clone() const {
    auto r = std::make_unique<XmlNode>();
    r->name = name;
    r->attrs = attrs;
    r->line = line;
    for (auto& c : children) {
        auto cc = c->clone();
        cc->parent = r.get();
        r->children.push_back(std::move(cc));
    }
    return r;
}

XmlNode* XmlNode::prepend_copy(const XmlNode& src) {
    auto c = src.clone();
    c->parent = this;
    XmlNode* p = c.get();
    children.insert(children.begin(), std::move(c));
    return p;
}

XmlNode* XmlNode::append_child(const std::string& n) {
    auto c = std::make_unique<XmlNode>();
    c->name = n;
    c->parent = this;
    XmlNode* p = c.get();
    children.push_back(std::move(c));
    return p;
}

void XmlNode::set(const std::string& n, const std::string& v) {
    for (auto& a : attrs)
        if (a.name == n) { a.value = v; return; }
    attrs.push_back({n, v});
}

namespace {

struct Parser {
    const std::string& s;
    size_t i = 0;
    int line = 1;
    explicit Parser(const std::string& t) : s(t) {}

    [[noreturn]] void fail(const std::string& what) const {
        throw std::runtime_error("XML parse error at line " + std::to_string(line) + ": " + what);
    }
    bool at_end() const { return i >= s.size(); }
    char peek(size_t k = 0) const { return i + k < s.size() ? s[i + k] : '\0'; }
    bool starts(const char* lit) const { return s.compare(i, std::char_traits<char>::length(lit), lit) == 0; }
    void adv(size_t k = 1) {
        for (size_t j = 0; j < k && i < s.size(); j++, i++)
            if (s[i] == '\n') line++;
    }
    void skip_space() {
        while (!at_end() && (peek() == ' ' || peek() == '\t' || peek() == '\r' || peek() == '\n')) adv();
    }
    void skip_until(const char* lit) {
        size_t p = s.find(lit, i);
        if (p == std::string::npos) fail(std::string("missing ") + lit);
        adv(p - i + std::char_traits<char>::length(lit));
    }
    static bool name_char(char c) {
        return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9') || c == '_' || c == '-' ||
               c == '.' || c == ':' || (unsigned char)c >= 0x80;
    }
    std::string name() {
        size_t b = i;
        while (!at_end() && name_char(peek())) adv();
        if (i == b) fail("expected a name");
        return s.substr(b, i - b);
    }
    static void put_utf8(std::string& o, unsigned long c) {
        if (c < 0x80) o += (char)c;
        else if (c < 0x800) { o += (char)(0xC0 | (c >> 6)); o += (char)(0x80 | (c & 0x3F)); }
        else if (c < 0x10000) {
            o += (char)(0xE0 | (c >> 12)); o += (char)(0x80 | ((c >> 6) & 0x3F)); o += (char)(0x80 | (c & 0x3F));
        } else {
            o += (char)(0xF0 | (c >> 18)); o += (char)(0x80 | ((c >> 12) & 0x3F));
            o += (char)(0x80 | ((c >> 6) & 0x3F)); o += (char)(0x80 | (c & 0x3F));
        }
    }
    // text up to `end` with entities decoded; attribute values turn
    // whitespace into spaces
    std::string decode(size_t end, bool attribute) {
        std::string o;
        while (i < end) {
            char c = peek();
            if (c == '&') {
                size_t semi = s.find(';', i);
                if (semi == std::string::npos || semi > end) { o += c; adv(); continue; }
                std::string ent = s.substr(i + 1, semi - i - 1);
                if (ent == "lt") o += '<';
                else if (ent == "gt") o += '>';
                else if (ent == "amp") o += '&';
                else if (ent == "quot") o += '"';
                else if (ent == "apos") o += '\'';
                else if (!ent.empty() && ent[0] == '#') {
                    unsigned long v = (ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X'))
                                          ? std::strtoul(ent.c_str() + 2, nullptr, 16)
                                          : std::strtoul(ent.c_str() + 1, nullptr, 10);
                    put_utf8(o, v);
                } else {
                    o += s.substr(i, semi - i + 1);  // unknown entity kept as written
                }
                adv(semi - i + 1);
                continue;
            }
            if (c == '\r') {  // end-of-line normalisation: CR LF and CR -> LF
                adv();
                if (peek() == '\n' && i < end) adv();
                o += attribute ? ' ' : '\n';
                continue;
            }
            if (attribute && (c == '\t' || c == '\n')) c = ' ';
            o += c;
            adv();
        }
        return o;
    }
    void misc() {
        for (;;) {
            skip_space();
            if (starts("<?")) skip_until("?>");
            else if (starts("<!--")) skip_until("-->");
            else if (starts("<!DOCTYPE")) skip_until(">");
            else return;
        }
    }
    std::unique_ptr<XmlNode> element(XmlNode* parent) {
        if (peek() != '<') fail("expected '<'");
        adv();
        auto n = std::make_unique<XmlNode>();
        n->parent = parent;
        n->line = line;
        n->name = name();
        for (;;) {
            skip_space();
            if (starts("/>")) { adv(2); return n; }
            if (peek() == '>') { adv(); break; }
            std::string an = name();
            skip_space();
            if (peek() != '=') fail("expected '=' after attribute " + an);
            adv();
            skip_space();
            char q = peek();
            if (q != '"' && q != '\'') fail("expected a quoted value for attribute " + an);
            adv();
            size_t e = s.find(q, i);
            if (e == std::string::npos) fail("unterminated attribute " + an);
            std::string v = decode(e, true);
            adv();
            n->attrs.push_back({an, v});
        }
        // content
        for (;;) {
            if (at_end()) fail("unterminated element <" + n->name + ">");
            if (starts("</")) {
                adv(2);
                std::string en = name();
                if (en != n->name) fail("</" + en + "> closes <" + n->name + ">");
                skip_space();
                if (peek() != '>') fail("expected '>'");
                adv();
                return n;
            }
            if (starts("<!--")) { skip_until("-->"); continue; }
            if (starts("<![CDATA[")) { skip_until("]]>"); continue; }
            if (starts("<?")) { skip_until("?>"); continue; }
            if (peek() == '<') {
                n->children.push_back(element(n.get()));
                continue;
            }
            size_t e = s.find('<', i);
            if (e == std::string::npos) e = s.size();
            decode(e, false);  // text content: not used by the case format
        }
    }
};

}  // namespace

std::unique_ptr<XmlNode> xml_parse(const std::string& text) {
    Parser p(text);
    auto doc = std::make_unique<XmlNode>();
    if (p.starts("\xEF\xBB\xBF")) p.adv(3);  // UTF-8 byte order mark
    p.misc();
    while (!p.at_end()) {
        doc->children.push_back(p.element(doc.get()));
        p.misc();
    }
    if (doc->children.empty()) p.fail("no root element");
    return doc;
}

std::unique_ptr<XmlNode> xml_load_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return xml_parse(ss.str());
}

namespace {
void write_node(const XmlNode& n, int depth, std::string& out) {
    out.append((size_t)depth, '\t');
    out += "<" + n.name;
    for (const XmlAttr& a : n.attrs) {
        out += " " + a.name + "=\"";
        for (char c : a.value) {
            switch (c) {
            case '&': out += "&amp;"; break;
            case '<': out += "&lt;"; break;
            case '>': out += "&gt;"; break;
            case '"': out += "&quot;"; break;
            default: out += c;
            }
        }
        out += "\"";
    }
    if (n.children.empty()) {
        out += " />\n";
        return;
    }
    out += ">\n";
    for (const auto& c : n.children) write_node(*c, depth + 1, out);
    out.append((size_t)depth, '\t');
    out += "</" + n.name + ">\n";
}
}  // namespace

std::string xml_write(const XmlNode& doc) {
    std::string out = "<?xml version=\"1.0\"?>\n";
    for (const auto& c : doc.children) write_node(*c, 0, out);
    return out;
}

int xml_as_int(const std::string& str) {
    const char* s = str.c_str();
    while (*s == ' ' || *s == '\t' || *s == '\r' || *s == '\n') s++;
    const bool neg = *s == '-';
    if (*s == '+' || *s == '-') s++;
    unsigned long long r = 0;
    bool overflow = false;
    if (s[0] == '0' && (s[1] | ' ') == 'x') {
        s += 2;
        for (;; s++) {
            int d;
            if ((unsigned)(*s - '0') < 10) d = *s - '0';
            else if ((unsigned)((*s | ' ') - 'a') < 6) d = (*s | ' ') - 'a' + 10;
            else break;
            r = r * 16 + d;
            if (r > 0xFFFFFFFFull) overflow = true;
        }
    } else {
        for (; (unsigned)(*s - '0') < 10; s++) {
            r = r * 10 + (*s - '0');
            if (r > 0xFFFFFFFFull) overflow = true;
        }
    }
    if (neg) return (overflow || r > 0x80000000ull) ? INT_MIN : (int)(0 - (long long)r);
    return (overflow || r > (unsigned long long)INT_MAX) ? INT_MAX : (int)r;
}

double xml_as_double(const std::string& s) { return std::strtod(s.c_str(), nullptr); }

bool xml_as_bool(const std::string& s) {
    const char c = s.empty() ? '\0' : s[0];
    return c == '1' || c == 't' || c == 'T' || c == 'y' || c == 'Y';
}

}  // namespace tclbx
