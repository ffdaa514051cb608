// NCF: the nested "name = value" / "name = {" text format that .scn scene files
// are written in. Behaviour follows lib/ncf/ncf.cc:89-271 and lib/ncf/util.cc
// (comment stripping, %include, <symbol> expansion against the parsing root's
// properties, std::map (alphabetical) ordering of groups and properties, and
// the same error codes and line numbers).
#pragma once
#include <map>
#include <memory>
#include <string>

namespace xt {
namespace ncf {

// Same numbering as lib/ncf/ncf.h:17-27.
enum Error {
    OK = 0,
    NOT_ASCII_FILE = 1,
    IO = 2,
    SYNTAX = 3,
    UNBALANCED_GROUPS = 4,
    AMBIGUOUS_GROUP = 5,
    AMBIGUOUS_PROPERTY = 6,
    EXTERNAL_FILE = 7,
};
const char *error_message(int code);

class Node {
public:
    // Parses `path` into this node (which acts as the root for symbol
    // expansion). Returns an Error code; *line receives the failing line.
    int parse(const std::string &path, size_t *line = nullptr);

    bool has_property(const std::string &name) const { return props_.count(name) != 0; }
    bool has_group(const std::string &name) const { return groups_.count(name) != 0; }
    const char *property(const std::string &name) const;  // nullptr if absent
    void set_property(const std::string &name, const std::string &value) { props_[name] = value; }
    // Like lib/ncf/ncf.cc:305-319, a missing group is created empty.
    Node *group(const std::string &name);
    const Node *find_group(const std::string &name) const;

    size_t count_groups() const { return groups_.size(); }
    size_t count_properties() const { return props_.size(); }
    Node *group_at(size_t i);                       // alphabetical order
    const std::string &group_name_at(size_t i) const;
    const std::string &property_name_at(size_t i) const;
    const std::string &property_at(size_t i) const;
    const std::string &name() const { return name_; }

private:
    int parse_into(const std::string &path, size_t *line);
    void expand(std::string &s) const;

    std::string name_;
    std::string path_;
    std::map<std::string, std::string> props_;
    std::map<std::string, std::unique_ptr<Node>> groups_;
};

}  // namespace ncf
}  // namespace xt
