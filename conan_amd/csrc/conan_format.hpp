// conan_format.hpp -- the .conan file format and conan's option/parameter
// system, for the host CLI.
//
// File layout (src/base/misc.hh:27-78, src/base/array.hh:75-105,
// src/base/header.cc:30-74, src/base/history.cc:25-69): a sequence of blocks,
// each framed by its byte count as uint64 before and after.  The first block
// is the parameter header ("key=value\n" lines sorted by key), the second the
// history ("time:command\n" lines sorted by time), then pairs of a record
// header (dtype=..., name=...) and the raw array (axis 0 fastest; particle
// records are AoS of dim doubles, dtype "3f8" / "2f8").
#pragma once
#include <cstdint>
#include <cstring>
#include <ctime>
#include <iosfwd>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace conan {

// Errors carry a message that main() prints as "[error] <msg>" (src/base/main.cc:101-116).
struct Failure {
    std::string msg;
};

void write_block(std::ostream& f, const void* data, uint64_t bytes);
void write_block_32(std::ostream& f, const void* data, uint32_t bytes);
std::vector<char> read_block(std::istream& f);
void skip_block(std::istream& f);

// ostream formatting of a value, as Misc::format / System::to_string do.
template <typename T>
std::string fmt(const T& v) {
    std::ostringstream s;
    s << v;
    return s.str();
}

template <typename T>
T parse(const std::string& s) {
    std::istringstream in(s);
    T v{};
    in >> v;
    return v;
}

// One command-line option (src/base/argv.hh:21-41).
struct Option {
    enum Flags { CHECK = 1, VALUED = 2 };
    int flags;
    std::string short_name, long_name, value, description;
    bool checked() const { return (flags & CHECK) != 0; }
    bool valued() const { return (flags & VALUED) != 0; }
};

class Argv {
  public:
    std::vector<Option> opts;
    std::string command_line;
    explicit Argv(std::vector<Option> o) : opts(std::move(o)) {}
    // Parses argv[n..]; returns the index of the first non-option argument.
    int parse(int argc, char** argv, int n = 1);
    std::string operator[](const std::string& name) const;
    template <typename T>
    T get(const std::string& name) const {
        return conan::parse<T>((*this)[name]);
    }
    void print(std::ostream& out) const;
};

// read_arguments (src/base/argv.hh:119-142): parse errors print
// "ERROR: <msg>" and exit.
Argv read_arguments(int argc, char** argv, std::vector<Option> opts);

class Header : public std::map<std::string, std::string> {
  public:
    Header() = default;
    explicit Header(std::istream& f) { from_file(f); }
    void from_file(std::istream& f);
    void to_file(std::ostream& f) const;
    // "# header", "# key=value" per entry, blank line (header.cc:44-52)
    void to_txt_file(std::ostream& f) const;
    bool has(const std::string& k) const { return find(k) != end(); }
    template <typename T>
    T get(const std::string& name) const {
        auto i = find(name);
        if (i == end()) throw Failure{"variable [" + name + "] not found in config."};
        return parse<T>(i->second);
    }
    // Header << Argv: every option, defaults included, overwrites (header.cc:22-28)
    Header& operator<<(const Argv& C);
};

class History : public std::multimap<time_t, std::string> {
  public:
    History() = default;
    explicit History(std::istream& f) { from_file(f); }
    void from_file(std::istream& f);
    void to_file(std::ostream& f) const;
    // "# history", "# time:command" per entry, blank line (history.cc:40-48)
    void to_txt_file(std::ostream& f) const;
    void update(const std::string& cmd) { insert({time(nullptr), cmd}); }
    History& operator<<(const Argv& C) {
        update(C.command_line);
        return *this;
    }
};

// save_to_file / load_from_file (src/base/array.hh:75-105)
void save_record(std::ostream& f, const std::string& name, const std::string& dtype, const void* data,
                 uint64_t bytes);
std::vector<char> load_record(std::istream& f, const std::string& name);
template <typename T>
std::vector<T> load_array(std::istream& f, const std::string& name) {
    std::vector<char> raw = load_record(f, name);
    std::vector<T> v(raw.size() / sizeof(T));
    if (!v.empty()) std::memcpy(v.data(), raw.data(), v.size() * sizeof(T));
    return v;
}

// yymmdd of today (src/base/date.cc:11-21)
std::string date_string();

}  // namespace conan
