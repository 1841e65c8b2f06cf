// A small XML reader for the two XML dialects the reconstruction reads: the
// parameter files (<a><b value="..."/></a> gives key "a.b") and the
// metadata.xml of an image-stack directory.  Both are written without a single
// root element, so a document here is a forest of top-level elements, as
// boost::property_tree::read_xml accepts them (XMLParameterSource.cpp:88-91,
// ImageStackDirectoryDatasource.cpp:181-185).
#pragma once

#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace etn {

struct XmlNode {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attributes;   // document order
    std::vector<std::unique_ptr<XmlNode>> children;                // document order

    // First child with this name (boost ptree get_child semantics), or null.
    const XmlNode* child(const std::string& n) const;
    // Attribute value or null.
    const std::string* attribute(const std::string& n) const;
};

// Parses text into a forest of top-level elements.  Declarations, comments,
// processing instructions and DOCTYPE are skipped; text content is ignored
// (the parameter files carry every value in an attribute).  Throws
// etn::Error(ETN_ERROR_INVALID_ARGUMENT) with the line number on malformed input.
std::vector<std::unique_ptr<XmlNode>> parse_xml(const std::string& text, const std::string& source_name);
std::vector<std::unique_ptr<XmlNode>> parse_xml_file(const std::string& path);

// Escapes &, <, >, " for an attribute value.
std::string xml_escape(const std::string& s);

}  // namespace etn
