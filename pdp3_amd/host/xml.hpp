// Minimal XML reader for parameters.xml (the reference reads it with
// tinyxml2 in COLLAPSE_WHITESPACE mode, src/parameters.cpp:9).
//
// Supported: the <?xml ...?> declaration, comments, elements with quoted
// attributes, self-closing elements, character data with the five predefined
// entities and numeric character references.  Text is whitespace-collapsed
// (leading/trailing blanks dropped, inner runs folded to one space), as
// tinyxml2 does in that mode.  Not supported: DTDs, CDATA, processing
// instructions other than the declaration.
#pragma once

#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace pdp3 {
namespace xml {

struct Element {
  std::string name;
  std::vector<std::pair<std::string, std::string>> attributes;
  std::string text;  // concatenated, whitespace-collapsed character data
  std::vector<std::unique_ptr<Element>> children;

  // first child with the given name, or nullptr (tinyxml2 FirstChildElement)
  const Element *first_child(const char *child_name) const;
  // all children with the given name, in document order
  std::vector<const Element *> children_named(const char *child_name) const;
  // attribute value, or nullptr when absent
  const char *attribute(const char *attr) const;
};

struct Document {
  std::unique_ptr<Element> root;
};

// parses a whole document; throws std::runtime_error with line information
Document parse(const std::string &text);
// reads and parses a file; throws std::runtime_error when unreadable
Document load_file(const std::string &path);

}  // namespace xml
}  // namespace pdp3
