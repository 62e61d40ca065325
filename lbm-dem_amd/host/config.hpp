// Reader for the reference's getpot-format configuration files and its command-line overrides.
//
// File: one `name = value` per line, `#` starts a comment, values may be quoted, `[section]`
// lines prefix the following names with `section/`. Command line: `-name value` overrides the
// file (the rule of the reference's BASE_PARSE_CFG, macros.h:14-22); `-c file` and `-d dir`
// select the configuration and the work directory (main.cpp:70-116).
#pragma once

#include <map>
#include <string>
#include <vector>

namespace lbdem {

class ConfigFile {
public:
    ConfigFile() = default;
    explicit ConfigFile(const std::string& path);
    bool ok() const { return ok_; }
    bool have_variable(const std::string& name) const { return vars_.count(name) != 0; }
    std::string raw(const std::string& name) const;
    double operator()(const std::string& name, double def) const;
    int operator()(const std::string& name, int def) const;
    std::string operator()(const std::string& name, const char* def) const;
    const std::map<std::string, std::string>& variables() const { return vars_; }

private:
    std::map<std::string, std::string> vars_;
    bool ok_ = false;
};

class CommandLine {
public:
    CommandLine(int argc, char** argv);
    explicit CommandLine(std::vector<std::string> args) : args_(std::move(args)) {}
    // true if `flag` is present; the following next() reads the argument after it
    bool search(const std::string& flag);
    std::string next(const std::string& def);
    std::string next(const char* def) { return next(std::string(def)); }
    double next(double def);
    int next(int def);

private:
    std::vector<std::string> args_;
    int cursor_ = -1;
};

// typed value conversions shared by both readers (getpot rules: "true"/"false" read as 1/0 for
// integers; a value that does not parse yields the default)
double to_double(const std::string& s, double def);
int to_int(const std::string& s, int def);

// value of `name`: the configuration file's, replaced by `-name value` on the command line; a
// name found in neither prints the reference's warning and yields `def`
double parse_value(const ConfigFile& cfg, CommandLine& cl, const std::string& name, double def);
int parse_value(const ConfigFile& cfg, CommandLine& cl, const std::string& name, int def);
std::string parse_value(const ConfigFile& cfg, CommandLine& cl, const std::string& name, const char* def);

}  // namespace lbdem
