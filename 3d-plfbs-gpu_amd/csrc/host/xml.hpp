// Minimal XML element-tree reader for case files: elements, text content,
// comments, declarations and attributes (attributes are skipped).  The case
// schema (src/preprocessor.cpp) uses nested elements with text only.
#pragma once
#include <memory>
#include <string>
#include <vector>

namespace plbfs {

struct XmlNode {
    std::string name;
    std::string text;
    std::vector<std::unique_ptr<XmlNode>> children;

    const XmlNode *child(const std::string &n) const {
        for (auto &c : children)
            if (c->name == n) return c.get();
        return nullptr;
    }
    // path like "rho/gradient/x"
    const XmlNode *path(const std::string &p) const {
        const XmlNode *cur = this;
        size_t s = 0;
        while (cur && s <= p.size()) {
            size_t e = p.find('/', s);
            if (e == std::string::npos) e = p.size();
            cur = cur->child(p.substr(s, e - s));
            s = e + 1;
            if (e == p.size()) break;
        }
        return cur;
    }
};

// Parses a document; returns the root element or nullptr with err set.
std::unique_ptr<XmlNode> xml_parse(const std::string &text, std::string &err);

}  // namespace plbfs
