// xml.h -- a small XML DOM for TCLB case files.
//
// TCLB reads its case files with pugixml (src/main.cpp:246, parse_default |
// parse_comments) and walks elements and attributes in document order.  The
// engine needs the same view: element names, attributes in order, children in
// order, parents, and the lookups the reference uses
// (child, find_child_by_attribute, prepend_copy).  Text content is kept but
// nothing in the case format reads it.  Comments, the XML declaration,
// processing instructions and DOCTYPE are skipped; the five predefined
// entities and numeric character references are decoded; in attribute
// values tab, CR and LF become spaces (pugixml's parse_wconv_attribute).
#pragma once
#include <memory>
#include <string>
#include <vector>

namespace tclbx {

struct XmlAttr {
    std::string name, value;
};

struct XmlNode {
    std::string name;
    std::vector<XmlAttr> attrs;
    std::vector<std::unique_ptr<XmlNode>> children;
    XmlNode* parent = nullptr;
    int line = 0;

    const XmlAttr* attr(const std::string& n) const;
    // value of attribute n, or `def` when it is absent
    std::string get(const std::string& n, const std::string& def = "") const;
    bool has(const std::string& n) const { return attr(n) != nullptr; }
    // first child element named n (nullptr if none)
    XmlNode* child(const std::string& n) const;
    // first child element named `elem` whose attribute `a` equals `v`
    XmlNode* find_child_by_attribute(const std::string& elem, const std::string& a, const std::string& v) const;
    // deep copy of `src` inserted as the first child; returns it
    XmlNode* prepend_copy(const XmlNode& src);
    XmlNode* append_child(const std::string& n);
    void set(const std::string& n, const std::string& v);
    std::unique_ptr<XmlNode> clone() const;
};

// Parses a whole document; the returned node is the document (empty name)
// whose children are the top-level elements.  Throws std::runtime_error
// with the line number on malformed input.
std::unique_ptr<XmlNode> xml_parse(const std::string& text);
std::unique_ptr<XmlNode> xml_load_file(const std::string& path);
// The document as pugixml's save_file writes it by default: the declaration
// <?xml version="1.0"?>, one element per line indented by tabs, empty
// elements as <name a="v" />, and &, <, >, " escaped in attribute values.
std::string xml_write(const XmlNode& doc);

// The value of a present attribute as pugixml's xml_attribute reads it:
// as_int skips leading space, takes an optional sign and decimal or 0x-hex
// digits, stops at the first other character (0 when there is none) and
// clamps to the int range; as_double is strtod (0 when nothing parses);
// as_bool is true when the first character is 1, t, T, y or Y.
int xml_as_int(const std::string& s);
double xml_as_double(const std::string& s);
bool xml_as_bool(const std::string& s);

}  // namespace tclbx
