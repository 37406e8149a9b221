// xml_reader.h -- single-pass streaming reader for the OpenMM-style system and
// state XML files.
//
// The whole file is read once; every element start is reported with the path
// of element names from the root and its attributes.  Text content, comments,
// processing instructions and DOCTYPE declarations are skipped.  Attribute
// values have the five predefined entities and numeric character references
// decoded.  A malformed document raises std::runtime_error with the byte
// offset of the problem.
#pragma once

#include <functional>
#include <string>
#include <utility>
#include <vector>

namespace nexusmd {

struct XmlAttrs {
    std::vector<std::pair<std::string, std::string>> items;
    // value of attribute `name`, or nullptr when it is absent
    const std::string* find(const char* name) const;
    // numeric readers; an absent or unparsable attribute reads as 0
    double as_double(const char* name) const;
    int as_int(const char* name) const;
    std::string as_string(const char* name) const;
};

// path: element names from the document root down to this element (inclusive)
using XmlStartFn = std::function<void(const std::vector<std::string>& path, const XmlAttrs& attrs)>;

// Reads `filename` and calls on_start for every element in document order.
// Returns false (and sets *error) when the file cannot be opened.
bool xml_scan_file(const std::string& filename, const XmlStartFn& on_start, std::string* error);

// Same over an in-memory document.
void xml_scan(const char* data, size_t size, const XmlStartFn& on_start);

}  // namespace nexusmd
