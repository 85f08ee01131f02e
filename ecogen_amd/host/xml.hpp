// Minimal XML reader for ECOGEN case files: elements, attributes, text,
// comments, processing instructions and the five predefined entities.  Enough
// for the reference's main/mesh/model/initialConditions and libEOS files; no
// DTDs, namespaces or CDATA.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace ecg_host {

struct XmlError : std::runtime_error {
  explicit XmlError(const std::string& m) : std::runtime_error(m) {}
};

struct XmlElement {
  std::string name;
  std::vector<std::pair<std::string, std::string>> attrs;
  std::vector<std::unique_ptr<XmlElement>> children;
  std::string text;   // concatenated character data directly inside the element

  // First child element with this name (nullptr if none).
  const XmlElement* child(const std::string& n) const;
  // All child elements with this name, in document order.
  std::vector<const XmlElement*> children_named(const std::string& n) const;
  // Attribute value, or nullptr when absent.
  const std::string* attr(const std::string& n) const;
  // Trimmed text content.
  std::string trimmed_text() const;
};

// Parses a document and returns its root element; throws XmlError with the
// file name and line of the first syntax error.
std::unique_ptr<XmlElement> parse_xml(const std::string& text, const std::string& where);
std::unique_ptr<XmlElement> load_xml(const std::string& path);

}  // namespace ecg_host
