#include "xml.hpp"

#include <cctype>
#include <fstream>
#include <sstream>

namespace ecg_host {

const XmlElement* XmlElement::child(const std::string& n) const {
  for (const auto& c : children)
    if (c->name == n) return c.get();
  return nullptr;
}

std::vector<const XmlElement*> XmlElement::children_named(const std::string& n) const {
  std::vector<const XmlElement*> out;
  for (const auto& c : children)
    if (c->name == n) out.push_back(c.get());
  return out;
}

const std::string* XmlElement::attr(const std::string& n) const {
  for (const auto& a : attrs)
    if (a.first == n) return &a.second;
  return nullptr;
}

std::string XmlElement::trimmed_text() const {
  size_t b = 0, e = text.size();
  while (b < e && std::isspace((unsigned char)text[b])) b++;
  while (e > b && std::isspace((unsigned char)text[e - 1])) e--;
  return text.substr(b, e - b);
}

namespace {

struct Parser {
  const std::string& s;
  const std::string& where;
  size_t i = 0;

  [[noreturn]] void fail(const std::string& msg) const {
    int line = 1;
    for (size_t k = 0; k < i && k < s.size(); k++)
      if (s[k] == '\n') line++;
    throw XmlError(where + ":" + std::to_string(line) + ": " + msg);
  }
  bool starts(const char* p) const { return s.compare(i, std::char_traits<char>::length(p), p) == 0; }
  void skip_ws() {
    while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
  }
  void skip_until(const char* end) {
    const size_t k = s.find(end, i);
    if (k == std::string::npos) fail(std::string("unterminated construct, expected '") + end + "'");
    i = k + std::char_traits<char>::length(end);
  }
  // Comments, processing instructions and declarations between elements.
  void skip_misc() {
    for (;;) {
      skip_ws();
      if (starts("<!--")) { i += 4; skip_until("-->"); }
      else if (starts("<?")) { i += 2; skip_until("?>"); }
      else if (starts("<!")) { i += 2; skip_until(">"); }
      else return;
    }
  }
  static bool name_char(char c) {
    return std::isalnum((unsigned char)c) || c == '_' || c == '-' || c == '.' || c == ':';
  }
  std::string name() {
    const size_t b = i;
    while (i < s.size() && name_char(s[i])) i++;
    if (i == b) fail("expected a name");
    return s.substr(b, i - b);
  }
  std::string decode(const std::string& v) {
    std::string o;
    for (size_t k = 0; k < v.size(); k++) {
      if (v[k] != '&') { o += v[k]; continue; }
      const size_t e = v.find(';', k);
      if (e == std::string::npos) fail("unterminated entity");
      const std::string ent = v.substr(k + 1, e - k - 1);
      if (ent == "lt") o += '<';
      else if (ent == "gt") o += '>';
      else if (ent == "amp") o += '&';
      else if (ent == "quot") o += '"';
      else if (ent == "apos") o += '\'';
      else fail("unknown entity &" + ent + ";");
      k = e;
    }
    return o;
  }
  std::unique_ptr<XmlElement> element() {
    if (!starts("<")) fail("expected '<'");
    i++;
    auto el = std::make_unique<XmlElement>();
    el->name = name();
    for (;;) {
      skip_ws();
      if (i >= s.size()) fail("unexpected end of file in tag <" + el->name + ">");
      if (starts("/>")) { i += 2; return el; }
      if (s[i] == '>') { i++; break; }
      std::string an = name();
      skip_ws();
      if (i >= s.size() || s[i] != '=') fail("expected '=' after attribute " + an);
      i++;
      skip_ws();
      if (i >= s.size() || (s[i] != '"' && s[i] != '\'')) fail("expected a quoted value for attribute " + an);
      const char q = s[i++];
      const size_t e = s.find(q, i);
      if (e == std::string::npos) fail("unterminated attribute value");
      el->attrs.emplace_back(an, decode(s.substr(i, e - i)));
      i = e + 1;
    }
    // content
    for (;;) {
      if (i >= s.size()) fail("unexpected end of file inside <" + el->name + ">");
      if (starts("<!--")) { i += 4; skip_until("-->"); continue; }
      if (starts("<?")) { i += 2; skip_until("?>"); continue; }
      if (starts("</")) {
        i += 2;
        const std::string close = name();
        if (close != el->name) fail("closing tag </" + close + "> does not match <" + el->name + ">");
        skip_ws();
        if (i >= s.size() || s[i] != '>') fail("expected '>'");
        i++;
        return el;
      }
      if (s[i] == '<') {
        el->children.push_back(element());
        continue;
      }
      const size_t e = s.find('<', i);
      el->text += decode(s.substr(i, (e == std::string::npos ? s.size() : e) - i));
      i = (e == std::string::npos) ? s.size() : e;
    }
  }
};

}  // namespace

std::unique_ptr<XmlElement> parse_xml(const std::string& text, const std::string& where) {
  Parser p{text, where};
  p.skip_misc();
  auto root = p.element();
  p.skip_misc();
  if (p.i < text.size()) p.fail("content after the root element");
  return root;
}

std::unique_ptr<XmlElement> load_xml(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw XmlError(path + ": cannot open file");
  std::stringstream ss;
  ss << f.rdbuf();
  return parse_xml(ss.str(), path);
}

}  // namespace ecg_host
