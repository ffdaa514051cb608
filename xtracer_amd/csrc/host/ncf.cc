#include "ncf.h"

#include <fstream>
#include <list>

namespace xt {
namespace ncf {

const char *error_message(int code) {
    static const char *msgs[] = {"ok",
                                 "not an ascii file",
                                 "io error",
                                 "syntax error",
                                 "unbalanced groups (missing bracket)",
                                 "group of the same name as the property already exists",
                                 "property of the same name as the group already exists",
                                 "error in external file"};
    return (code >= 0 && code <= 7) ? msgs[code] : "unknown error";
}

namespace {

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; }

std::string trimmed(const std::string &s) {
    size_t b = 0, e = s.size();
    while (b < e && is_space(s[b])) ++b;
    while (e > b && is_space(s[e - 1])) --e;
    return s.substr(b, e - b);
}

// Split at the first occurrence of c (lib/ncf/util.cc split()).
void split(const std::string &in, std::string &l, std::string &r, char c) {
    size_t p = in.find(c);
    if (p == std::string::npos) {
        l = in;
        r.clear();
    } else {
        l = in.substr(0, p);
        r = in.substr(p + 1);
    }
}

// Locate the innermost-leftmost "<...>" of a balanced string, remove it from
// src and return it (with the brackets); empty if there is none.
std::string take_pattern(std::string &src) {
    int open = 0, close = 0;
    for (char ch : src) {
        if (ch == '<') ++open;
        else if (ch == '>') ++close;
        if (close > open) return {};
    }
    if (open != close || open == 0) return {};
    size_t i2 = src.find('>');
    size_t i1 = src.rfind('<', i2);
    if (i1 == std::string::npos) return {};
    std::string t = src.substr(i1, i2 + 1 - i1);
    src = src.substr(0, i1) + src.substr(i2 + 1);
    return t;
}

std::string dir_of(const std::string &path) {
    size_t p = path.find_last_of('/');
    return p == std::string::npos ? std::string() : path.substr(0, p + 1);
}

}  // namespace

const char *Node::property(const std::string &name) const {
    auto it = props_.find(name);
    return it == props_.end() ? nullptr : it->second.c_str();
}

Node *Node::group(const std::string &name) {
    auto &slot = groups_[name];
    if (!slot) {
        slot.reset(new Node);
        slot->name_ = name;
    }
    return slot.get();
}

const Node *Node::find_group(const std::string &name) const {
    auto it = groups_.find(name);
    return it == groups_.end() ? nullptr : it->second.get();
}

Node *Node::group_at(size_t i) {
    auto it = groups_.begin();
    std::advance(it, i);
    return it->second.get();
}
const std::string &Node::group_name_at(size_t i) const {
    auto it = groups_.begin();
    std::advance(it, i);
    return it->first;
}
const std::string &Node::property_name_at(size_t i) const {
    auto it = props_.begin();
    std::advance(it, i);
    return it->first;
}
const std::string &Node::property_at(size_t i) const {
    auto it = props_.begin();
    std::advance(it, i);
    return it->second;
}

// Replace "<name>" by the value of this (root) node's property `name`; an
// unknown name is left in place (lib/ncf/ncf.cc:232-271).
void Node::expand(std::string &s) const {
    if (s.size() < 3) return;
    size_t lt = s.find('<');
    if (lt == std::string::npos || lt == s.rfind('>')) return;
    std::string src = s;
    for (;;) {
        std::string pat = take_pattern(src);
        if (pat.size() < 3) return;
        std::string key = pat.substr(1, pat.size() - 2);
        auto it = props_.find(key);
        if (it != props_.end()) {
            size_t at = s.find(pat);
            if (at != std::string::npos) s = s.substr(0, at) + it->second + s.substr(at + pat.size());
        }
    }
}

int Node::parse(const std::string &path, size_t *line) {
    if (line) *line = 0;
    path_ = path;
    return parse_into(path, line);
}

// %include paths resolve against path_ of the node the parse was started on;
// an include at that node's top level moves path_ to the included file, and an
// include inside a group hands the group its own path (lib/ncf/ncf.cc:138-152).
int Node::parse_into(const std::string &path, size_t *err_line) {
    {
        std::ifstream chk(path, std::ios::binary);
        if (!chk.good()) return IO;
        int c;
        while ((c = chk.get()) != EOF)
            if (c > 127) return NOT_ASCII_FILE;
    }
    std::ifstream in(path);
    if (!in.good()) return IO;

    std::list<Node *> stack{this};
    size_t lineno = 0;
    int depth = 0;
    std::string input;
    auto fail = [&](int code) {
        if (err_line) *err_line = lineno;
        return code;
    };
    while (std::getline(in, input)) {
        ++lineno;
        std::string body, comment;
        split(input, body, comment, '#');
        std::string ln = trimmed(body);
        if (ln.empty()) continue;
        if (ln.size() > 1 && ln.find('%') != std::string::npos) {
            std::string com, value;
            split(ln, com, value, ' ');
            com = trimmed(com);
            value = trimmed(value);
            if (com != "%include") return fail(SYNTAX);
            Node *target = stack.front();
            std::string inc = dir_of(path_) + value;
            target->path_ = inc;
            if (target->parse_into(inc, nullptr) != OK) return fail(EXTERNAL_FILE);
        } else if (ln.size() > 2 && ln.find('=') != std::string::npos) {
            std::string name, value;
            split(ln, name, value, '=');
            name = trimmed(name);
            value = trimmed(value);
            if (value == "{") {
                ++depth;
                Node *g = stack.front()->group(name);
                if (stack.front()->has_property(name)) return fail(AMBIGUOUS_GROUP);
                stack.push_front(g);
            } else {
                for (size_t k = 0; k < stack.size(); ++k) expand(value);
                if (stack.front()->has_group(name)) return fail(AMBIGUOUS_PROPERTY);
                stack.front()->props_[name] = value;
            }
        } else if (ln.size() == 1 && ln[0] == '}') {
            if (--depth < 0) return fail(UNBALANCED_GROUPS);
            stack.pop_front();
        } else {
            return fail(SYNTAX);
        }
    }
    if (depth) return fail(UNBALANCED_GROUPS);
    return OK;
}

}  // namespace ncf
}  // namespace xt
