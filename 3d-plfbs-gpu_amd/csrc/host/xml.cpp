// Minimal recursive-descent XML reader (see xml.hpp).
#include "xml.hpp"

#include <cctype>

namespace plbfs {

namespace {

// Case files nest a handful of levels; deeper documents are rejected rather
// than recursed into without bound.
constexpr int kMaxDepth = 64;

struct Parser {
    const std::string &s;
    size_t i = 0;
    std::string err;
    explicit Parser(const std::string &t) : s(t) {}

    bool eof() const { return i >= s.size(); }
    void skip_ws() {
        while (!eof() && std::isspace((unsigned char)s[i])) i++;
    }
    bool starts(const char *p) const { return s.compare(i, std::char_traits<char>::length(p), p) == 0; }
    // skips comments, declarations, processing instructions, doctype
    bool skip_misc() {
        for (;;) {
            skip_ws();
            if (starts("<!--")) {
                size_t e = s.find("-->", i + 4);
                if (e == std::string::npos) return fail("unterminated comment");
                i = e + 3;
            } else if (starts("<?")) {
                size_t e = s.find("?>", i + 2);
                if (e == std::string::npos) return fail("unterminated declaration");
                i = e + 2;
            } else if (starts("<!")) {
                size_t e = s.find('>', i + 2);
                if (e == std::string::npos) return fail("unterminated markup");
                i = e + 1;
            } else {
                return true;
            }
        }
    }
    bool fail(const std::string &m) {
        if (err.empty()) err = m + " at offset " + std::to_string(i);
        return false;
    }
    static std::string decode(const std::string &t) {
        std::string o;
        for (size_t k = 0; k < t.size(); k++) {
            if (t[k] == '&') {
                size_t e = t.find(';', k);
                if (e != std::string::npos) {
                    std::string ent = t.substr(k + 1, e - k - 1);
                    if (ent == "lt") o += '<';
                    else if (ent == "gt") o += '>';
                    else if (ent == "amp") o += '&';
                    else if (ent == "quot") o += '"';
                    else if (ent == "apos") o += '\'';
                    else o += t.substr(k, e - k + 1);
                    k = e;
                    continue;
                }
            }
            o += t[k];
        }
        return o;
    }
    static std::string trim(const std::string &t) {
        size_t a = 0, b = t.size();
        while (a < b && std::isspace((unsigned char)t[a])) a++;
        while (b > a && std::isspace((unsigned char)t[b - 1])) b--;
        return t.substr(a, b - a);
    }
    std::unique_ptr<XmlNode> element(int depth = 0) {
        if (eof() || s[i] != '<') {
            fail("expected element");
            return nullptr;
        }
        if (depth >= kMaxDepth) {
            fail("elements nested deeper than " + std::to_string(kMaxDepth));
            return nullptr;
        }
        i++;
        size_t n0 = i;
        while (!eof() && !std::isspace((unsigned char)s[i]) && s[i] != '>' && s[i] != '/') i++;
        auto node = std::make_unique<XmlNode>();
        node->name = s.substr(n0, i - n0);
        if (node->name.empty()) {
            fail("empty element name");
            return nullptr;
        }
        // attributes (skipped), honouring quotes
        while (!eof() && s[i] != '>' && !(s[i] == '/' && i + 1 < s.size() && s[i + 1] == '>')) {
            if (s[i] == '"' || s[i] == '\'') {
                char q = s[i++];
                while (!eof() && s[i] != q) i++;
            }
            i++;
        }
        if (eof()) {
            fail("unterminated start tag");
            return nullptr;
        }
        if (s[i] == '/') {
            i += 2;
            return node;
        }
        i++;  // '>'
        std::string text;
        for (;;) {
            if (eof()) {
                fail("unterminated element " + node->name);
                return nullptr;
            }
            if (starts("</")) {
                i += 2;
                size_t e = s.find('>', i);
                if (e == std::string::npos) {
                    fail("unterminated end tag");
                    return nullptr;
                }
                if (trim(s.substr(i, e - i)) != node->name) {
                    fail("mismatched end tag for " + node->name);
                    return nullptr;
                }
                i = e + 1;
                break;
            }
            if (starts("<!--") || starts("<?")) {
                if (!skip_misc()) return nullptr;
                continue;
            }
            if (starts("<![CDATA[")) {
                size_t e = s.find("]]>", i);
                if (e == std::string::npos) {
                    fail("unterminated CDATA");
                    return nullptr;
                }
                text += s.substr(i + 9, e - i - 9);
                i = e + 3;
                continue;
            }
            if (s[i] == '<') {
                auto c = element(depth + 1);
                if (!c) return nullptr;
                node->children.push_back(std::move(c));
                continue;
            }
            text += s[i++];
        }
        node->text = trim(decode(text));
        return node;
    }
};

}  // namespace

std::unique_ptr<XmlNode> xml_parse(const std::string &text, std::string &err) {
    Parser p(text);
    if (!p.skip_misc()) {
        err = p.err;
        return nullptr;
    }
    auto root = p.element();
    if (!root) err = p.err;
    return root;
}

}  // namespace plbfs
