#include "config.hpp"

#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>

namespace lbdem {

static std::string trim(const std::string& s) {
    const size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

ConfigFile::ConfigFile(const std::string& path) {
    std::ifstream in(path);
    if (!in) return;
    ok_ = true;
    std::string line, section;
    while (std::getline(in, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line = line.substr(0, hash);
        line = trim(line);
        if (line.empty()) continue;
        if (line.front() == '[' && line.back() == ']') {
            section = trim(line.substr(1, line.size() - 2));
            if (!section.empty() && section.back() != '/') section += '/';
            continue;
        }
        const size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        const std::string name = trim(line.substr(0, eq));
        std::string val = trim(line.substr(eq + 1));
        if (val.size() >= 2 && (val.front() == '"' || val.front() == '\'') && val.back() == val.front())
            val = val.substr(1, val.size() - 2);
        if (!name.empty()) vars_[section + name] = val;
    }
}

std::string ConfigFile::raw(const std::string& name) const {
    const auto it = vars_.find(name);
    return it == vars_.end() ? std::string() : it->second;
}

double to_double(const std::string& s, double def) {
    const std::string t = trim(s);
    if (t == "true") return 1.0;
    if (t == "false") return 0.0;
    char* end = nullptr;
    const double v = std::strtod(t.c_str(), &end);
    return (end && end != t.c_str()) ? v : def;
}

int to_int(const std::string& s, int def) {
    const std::string t = trim(s);
    if (t == "true") return 1;
    if (t == "false") return 0;
    char* end = nullptr;
    const long v = std::strtol(t.c_str(), &end, 10);
    if (!end || end == t.c_str()) return def;
    if (*end == '.' || *end == 'e' || *end == 'E') return int(to_double(t, double(def)));
    return int(v);
}

double ConfigFile::operator()(const std::string& name, double def) const {
    return have_variable(name) ? to_double(raw(name), def) : def;
}
int ConfigFile::operator()(const std::string& name, int def) const {
    return have_variable(name) ? to_int(raw(name), def) : def;
}
std::string ConfigFile::operator()(const std::string& name, const char* def) const {
    return have_variable(name) ? raw(name) : std::string(def);
}

CommandLine::CommandLine(int argc, char** argv) {
    for (int i = 0; i < argc; ++i) args_.emplace_back(argv[i]);
}

bool CommandLine::search(const std::string& flag) {
    for (size_t i = 1; i < args_.size(); ++i)
        if (args_[i] == flag) {
            cursor_ = int(i);
            return true;
        }
    cursor_ = -1;
    return false;
}

std::string CommandLine::next(const std::string& def) {
    if (cursor_ < 0 || cursor_ + 1 >= int(args_.size())) return def;
    return args_[++cursor_];
}
double CommandLine::next(double def) { return to_double(next(std::string()), def); }
int CommandLine::next(int def) { return to_int(next(std::string()), def); }

template <typename T>
static void warn_missing(const ConfigFile& cfg, CommandLine& cl, const std::string& name, const T& def) {
    if (!cfg.have_variable(name) && !cl.search("-" + name))
        std::cout << "WARNING: Parameter " << name << " is missing! Using default: " << def << std::endl;
}

double parse_value(const ConfigFile& cfg, CommandLine& cl, const std::string& name, double def) {
    warn_missing(cfg, cl, name, def);
    double v = cfg(name, def);
    if (cl.search("-" + name)) v = cl.next(v);
    return v;
}
int parse_value(const ConfigFile& cfg, CommandLine& cl, const std::string& name, int def) {
    warn_missing(cfg, cl, name, def);
    int v = cfg(name, def);
    if (cl.search("-" + name)) v = cl.next(v);
    return v;
}
std::string parse_value(const ConfigFile& cfg, CommandLine& cl, const std::string& name, const char* def) {
    warn_missing(cfg, cl, name, def);
    std::string v = cfg(name, def);
    if (cl.search("-" + name)) v = cl.next(v);
    return v;
}

}  // namespace lbdem
