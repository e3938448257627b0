// conan_format.cpp -- .conan blocks, parameter headers, history, records and
// the option parser of the conan CLI (see conan_format.hpp).
#include <cstring>

#include "conan_format.hpp"

#include <algorithm>
#include <iomanip>
#include <iostream>

namespace conan {

void write_block(std::ostream& f, const void* data, uint64_t bytes) {
    f.write(reinterpret_cast<const char*>(&bytes), sizeof bytes);
    if (bytes) f.write(reinterpret_cast<const char*>(data), (std::streamsize)bytes);
    f.write(reinterpret_cast<const char*>(&bytes), sizeof bytes);
}

void write_block_32(std::ostream& f, const void* data, uint32_t bytes) {
    f.write(reinterpret_cast<const char*>(&bytes), sizeof bytes);
    if (bytes) f.write(reinterpret_cast<const char*>(data), (std::streamsize)bytes);
    f.write(reinterpret_cast<const char*>(&bytes), sizeof bytes);
}

// src/base/misc.hh:27-42
std::vector<char> read_block(std::istream& f) {
    uint64_t b1 = 0, b2 = 0;
    f.read(reinterpret_cast<char*>(&b1), sizeof b1);
    if (f.fail()) throw Failure{"Could not read from file"};
    std::vector<char> v(b1);
    if (b1) f.read(v.data(), (std::streamsize)b1);
    f.read(reinterpret_cast<char*>(&b2), sizeof b2);
    if (b1 != b2) throw Failure{"The block seems to be corrupted"};
    return v;
}

// src/base/misc.hh:67-78
void skip_block(std::istream& f) {
    uint64_t b1 = 0, b2 = 0;
    f.read(reinterpret_cast<char*>(&b1), sizeof b1);
    if (f.fail()) throw Failure{"Could not read from file"};
    f.seekg((std::streamoff)b1, std::ios::cur);
    f.read(reinterpret_cast<char*>(&b2), sizeof b2);
    if (b1 != b2) throw Failure{"The block seems to be corrupted"};
}

static std::string first_line(const std::string& s) { return s.substr(0, s.find('\n')); }

// ------------------------------------------------------------ Header
// "name=value\n" lines; an existing key is not replaced (header.cc:46-60)
void Header::from_file(std::istream& f) {
    clear();
    std::vector<char> raw = read_block(f);
    std::string in(raw.begin(), raw.end());
    size_t pos = 0;
    while (in.length() > pos) {
        size_t eq = in.find('=', pos), nl = in.find('\n', pos);
        if (nl == std::string::npos) nl = in.length();
        if (eq == std::string::npos || eq > nl) eq = nl;
        std::string name = in.substr(pos, eq - pos);
        std::string value = eq < nl ? in.substr(eq + 1, nl - eq - 1) : std::string();
        insert(value_type(name, value));
        pos = nl + 1;
    }
}

void Header::to_file(std::ostream& f) const {
    std::ostringstream s;
    for (const auto& kv : *this) s << kv.first << "=" << first_line(kv.second) << "\n";
    const std::string out = s.str();
    write_block(f, out.data(), out.size());
}

void Header::to_txt_file(std::ostream& f) const {
    f << "# header\n";
    for (const auto& kv : *this) f << "# " << kv.first << "=" << first_line(kv.second) << std::endl;
    f << std::endl;
}

Header& Header::operator<<(const Argv& C) {
    for (const auto& o : C.opts) (*this)[o.long_name] = o.value;
    return *this;
}

// ------------------------------------------------------------ History
void History::from_file(std::istream& f) {
    std::vector<char> raw = read_block(f);
    std::string in(raw.begin(), raw.end());
    size_t pos = 0;
    while (in.length() > pos) {
        size_t c = in.find(':', pos), nl = in.find('\n', pos);
        if (nl == std::string::npos) nl = in.length();
        if (c == std::string::npos || c > nl) c = nl;
        time_t t = parse<time_t>(in.substr(pos, c - pos));
        insert(value_type(t, c < nl ? in.substr(c + 1, nl - c - 1) : std::string()));
        pos = nl + 1;
    }
}

void History::to_file(std::ostream& f) const {
    std::ostringstream s;
    for (const auto& kv : *this) s << kv.first << ":" << first_line(kv.second) << "\n";
    const std::string out = s.str();
    write_block(f, out.data(), out.size());
}

void History::to_txt_file(std::ostream& f) const {
    f << "# history\n";
    for (const auto& kv : *this) f << "# " << kv.first << ":" << first_line(kv.second) << std::endl;
    f << std::endl;
}

// ------------------------------------------------------------ records
void save_record(std::ostream& f, const std::string& name, const std::string& dtype, const void* data,
                 uint64_t bytes) {
    Header S;
    S["name"] = name;
    S["dtype"] = dtype;
    S.to_file(f);
    write_block(f, data, bytes);
}

// Scans record headers from the current position; only "name" is checked,
// so files written by the reference's Python tools (which say "type=") load
// too.  The stream is put back where it was (array.hh:84-105).
std::vector<char> load_record(std::istream& f, const std::string& name) {
    auto pos = f.tellg();
    while (f.good()) {
        Header S(f);
        if (S["name"] != name) {
            skip_block(f);
        } else {
            std::vector<char> data = read_block(f);
            f.clear();
            f.seekg(pos, std::ios::beg);
            return data;
        }
    }
    throw Failure{"Couldn't find record " + name + " in file."};
}

// ------------------------------------------------------------ Argv
// src/base/argv.cc:42-99
int Argv::parse(int argc, char** argv, int n) {
    command_line = argv[0];
    for (int i = 1; i < argc; ++i) command_line += std::string(" ") + argv[i];
    while (n != argc) {
        const std::string arg(argv[n]);
        if (arg.empty() || arg[0] != '-') return n;
        const bool lng = arg.size() > 1 && arg[1] == '-';
        const std::string name = lng ? arg.substr(2, 100) : arg.substr(1, 100);
        auto it = std::find_if(opts.begin(), opts.end(),
                               [&](const Option& o) { return lng ? o.long_name == name : o.short_name == name; });
        if (it == opts.end()) throw Failure{"Got the argument [" + arg + "], but no such option exists."};
        if (it->valued() && (argc - n) < 2)
            throw Failure{"Got the last argument [" + arg + "], but no value given."};
        if (it->valued()) {
            it->value = argv[n + 1];
            it->flags |= Option::CHECK;
            n += 2;
        } else {
            it->flags |= Option::CHECK;
            it->value = "1";
            n += 1;
        }
    }
    return n;
}

std::string Argv::operator[](const std::string& name) const {
    for (const auto& o : opts)
        if (o.long_name == name) return o.value;
    throw Failure{"Trying to find option [" + name + "], but no such option exists."};
}

// Help layout of operator<<(ostream, Option) (src/base/argv.cc:109-139):
// "        --long | -s (default: v) : words..." wrapped before 72 columns,
// continuation lines indented by 16 spaces.
void Argv::print(std::ostream& out) const {
    for (const auto& o : opts) {
        std::string line = "        --" + o.long_name;
        if (!o.short_name.empty()) line += " | -" + o.short_name;
        if (o.valued() && o.checked()) line += " (default: " + o.value + ")";
        if (o.valued() && !o.checked()) line += " (obligatory)";
        if (!o.valued()) line += " (flag)";
        line += " : ";
        std::vector<std::string> words;
        std::istringstream ws(o.description);
        for (std::string w; ws >> w;) words.push_back(w);
        size_t i = 0;
        while (i < words.size()) {
            do {
                line += words[i] + " ";
                ++i;
            } while (line.length() < 72 && i < words.size());
            out << line << "\n";
            line = std::string(16, ' ');
        }
        out << "\n";
    }
}

Argv read_arguments(int argc, char** argv, std::vector<Option> opts) {
    Argv C(std::move(opts));
    try {
        int n = C.parse(argc, argv);
        if (n != argc) throw Failure{std::string("Don't know how to handle [") + argv[n] + "]"};
    } catch (const Failure& e) {
        std::cout << "ERROR: " << e.msg << "\n";
        exit(-1);
    }
    return C;
}

std::string date_string() {
    time_t t = time(nullptr);
    struct tm* T = gmtime(&t);
    std::ostringstream s;
    s << std::setfill('0') << std::setw(2) << T->tm_year - 100 << std::setw(2) << T->tm_mon + 1 << std::setw(2)
      << T->tm_mday;
    return s.str();
}

}  // namespace conan
