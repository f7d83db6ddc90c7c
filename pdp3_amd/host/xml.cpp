#include "xml.hpp"

#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace pdp3 {
namespace xml {

const Element *Element::first_child(const char *child_name) const {
  for (const auto &c : children)
    if (c->name == child_name) return c.get();
  return nullptr;
}

std::vector<const Element *> Element::children_named(const char *child_name) const {
  std::vector<const Element *> out;
  for (const auto &c : children)
    if (c->name == child_name) out.push_back(c.get());
  return out;
}

const char *Element::attribute(const char *attr) const {
  for (const auto &a : attributes)
    if (a.first == attr) return a.second.c_str();
  return nullptr;
}

namespace {

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; }
bool is_name_char(char c) {
  return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9') || c == '_' || c == '-' ||
         c == '.' || c == ':';
}

void append_utf8(std::string &out, unsigned long cp) {
  if (cp < 0x80) {
    out += (char)cp;
  } else if (cp < 0x800) {
    out += (char)(0xC0 | (cp >> 6));
    out += (char)(0x80 | (cp & 0x3F));
  } else if (cp < 0x10000) {
    out += (char)(0xE0 | (cp >> 12));
    out += (char)(0x80 | ((cp >> 6) & 0x3F));
    out += (char)(0x80 | (cp & 0x3F));
  } else {
    out += (char)(0xF0 | (cp >> 18));
    out += (char)(0x80 | ((cp >> 12) & 0x3F));
    out += (char)(0x80 | ((cp >> 6) & 0x3F));
    out += (char)(0x80 | (cp & 0x3F));
  }
}

class Parser {
 public:
  explicit Parser(const std::string &s) : s_(s) {}

  Document run() {
    Document doc;
    skip_misc();
    if (pos_ >= s_.size() || s_[pos_] != '<') fail("no root element");
    doc.root = element();
    skip_misc();
    if (pos_ != s_.size()) fail("content after the root element");
    return doc;
  }

 private:
  const std::string &s_;
  size_t pos_ = 0;

  [[noreturn]] void fail(const std::string &msg) const {
    int line = 1;
    for (size_t i = 0; i < pos_ && i < s_.size(); i++)
      if (s_[i] == '\n') line++;
    throw std::runtime_error("XML parse error at line " + std::to_string(line) + ": " + msg);
  }
  bool starts(const char *lit) const { return s_.compare(pos_, std::strlen(lit), lit) == 0; }
  void skip_ws() {
    while (pos_ < s_.size() && is_space(s_[pos_])) pos_++;
  }
  void skip_until(const char *lit) {
    size_t e = s_.find(lit, pos_);
    if (e == std::string::npos) fail(std::string("unterminated construct, expected ") + lit);
    pos_ = e + std::strlen(lit);
  }
  // whitespace, comments and the declaration between top-level items
  void skip_misc() {
    for (;;) {
      skip_ws();
      if (starts("<!--")) {
        skip_until("-->");
      } else if (starts("<?")) {
        skip_until("?>");
      } else if (starts("<!")) {
        fail("DTD / CDATA sections are not supported");
      } else {
        return;
      }
    }
  }
  std::string name() {
    size_t b = pos_;
    while (pos_ < s_.size() && is_name_char(s_[pos_])) pos_++;
    if (b == pos_) fail("expected a name");
    return s_.substr(b, pos_ - b);
  }
  std::string decode(const std::string &raw) {
    std::string out;
    for (size_t i = 0; i < raw.size(); i++) {
      if (raw[i] != '&') {
        out += raw[i];
        continue;
      }
      size_t semi = raw.find(';', i);
      if (semi == std::string::npos) fail("unterminated entity");
      std::string ent = raw.substr(i + 1, semi - i - 1);
      if (ent == "lt") out += '<';
      else if (ent == "gt") out += '>';
      else if (ent == "amp") out += '&';
      else if (ent == "quot") out += '"';
      else if (ent == "apos") out += '\'';
      else if (!ent.empty() && ent[0] == '#') {
        unsigned long cp = ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X') ? std::stoul(ent.substr(2), nullptr, 16)
                                                                               : std::stoul(ent.substr(1), nullptr, 10);
        append_utf8(out, cp);
      } else {
        fail("unknown entity &" + ent + ";");
      }
      i = semi;
    }
    return out;
  }
  static std::string collapse(const std::string &t) {
    std::string out;
    bool pending_space = false;
    for (char c : t) {
      if (is_space(c)) {
        pending_space = !out.empty();
      } else {
        if (pending_space) out += ' ';
        pending_space = false;
        out += c;
      }
    }
    return out;
  }
  std::unique_ptr<Element> element() {
    pos_++;  // '<'
    auto el = std::make_unique<Element>();
    el->name = name();
    for (;;) {
      skip_ws();
      if (pos_ >= s_.size()) fail("unterminated start tag <" + el->name + ">");
      if (starts("/>")) {
        pos_ += 2;
        return el;
      }
      if (s_[pos_] == '>') {
        pos_++;
        break;
      }
      std::string an = name();
      skip_ws();
      if (pos_ >= s_.size() || s_[pos_] != '=') fail("expected '=' after attribute " + an);
      pos_++;
      skip_ws();
      if (pos_ >= s_.size() || (s_[pos_] != '"' && s_[pos_] != '\'')) fail("expected a quoted value for " + an);
      char q = s_[pos_++];
      size_t e = s_.find(q, pos_);
      if (e == std::string::npos) fail("unterminated attribute value");
      el->attributes.emplace_back(an, decode(s_.substr(pos_, e - pos_)));
      pos_ = e + 1;
    }
    std::string raw_text;
    for (;;) {
      if (pos_ >= s_.size()) fail("missing </" + el->name + ">");
      if (starts("<!--")) {
        skip_until("-->");
      } else if (starts("</")) {
        pos_ += 2;
        std::string en = name();
        if (en != el->name) fail("mismatched </" + en + "> for <" + el->name + ">");
        skip_ws();
        if (pos_ >= s_.size() || s_[pos_] != '>') fail("expected '>'");
        pos_++;
        break;
      } else if (starts("<!") || starts("<?")) {
        fail("unsupported markup inside <" + el->name + ">");
      } else if (s_[pos_] == '<') {
        el->children.push_back(element());
      } else {
        size_t e = s_.find('<', pos_);
        if (e == std::string::npos) e = s_.size();
        raw_text += s_.substr(pos_, e - pos_);
        pos_ = e;
      }
    }
    el->text = collapse(decode(raw_text));
    return el;
  }
};

}  // namespace

Document parse(const std::string &text) { return Parser(text).run(); }

Document load_file(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("can not read configuration file ``" + path + "``");
  std::stringstream ss;
  ss << f.rdbuf();
  return parse(ss.str());
}

}  // namespace xml
}  // namespace pdp3
