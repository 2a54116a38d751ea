// Logging, limits and command-line helpers of the graph_one C++ API:
//   LOG_INFO/LOG_WARNING/LOG_DEBUG/LOG_ERROR (exits with status 1) and printx
//     (reference: include/GraphOne/utils/log.hpp:30-63),
//   invalid_vertex / is_vertex_valid (include/GraphOne/utils/limits.hpp:54-73),
//   a small command-line parser with the subset of the CLI11 interface the apps use (CLI::App,
//   add_option(...)->required(), add_flag, CLI11_PARSE) and add_common_args (include/GraphOne/utils/cli.hpp:18-32).
#pragma once

#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <limits>
#include <memory>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>

#include "GraphOne/type.hpp"

namespace graph_one {

namespace utils {

template <typename T, typename = void>
struct HasToStrMethod : std::false_type {};
template <typename T>
struct HasToStrMethod<T, std::void_t<decltype(std::declval<const T&>().ToString())>> : std::true_type {};

template <typename T>
std::string NumToString(const T& v) {
    std::ostringstream os;
    os << v;
    return os.str();
}

inline std::string ToString(bool b) { return b ? "true" : "false"; }
inline std::string ToString(arch_t a) { return a == arch_t::cpu ? "cpu" : "hip"; }
template <typename T>
std::enable_if_t<std::is_arithmetic_v<T>, std::string> ToString(const T& v) {
    return NumToString(v);
}

inline bool StrEndWith(const std::string& s, const std::string& suffix) {
    return s.size() >= suffix.size() && s.compare(s.size() - suffix.size(), suffix.size(), suffix) == 0;
}
inline bool StrStartWith(const std::string& s, const std::string& prefix) {
    return !prefix.empty() && s.compare(0, prefix.size(), prefix) == 0;
}
inline std::string StrStrip(const std::string& s) {
    const char* ws = " \t\r\n";
    const size_t b = s.find_first_not_of(ws);
    if (b == std::string::npos) return "";
    const size_t e = s.find_last_not_of(ws);
    return s.substr(b, e - b + 1);
}

// The invalid vertex id: the largest value for ids starting at 0, 0 for ids starting at 1.
template <vstart_t v_start, typename vertex_t>
constexpr vertex_t invalid_vertex() {
    if constexpr (v_start == vstart_t::FROM_0_TO_0) return std::numeric_limits<vertex_t>::max();
    else return vertex_t(0);
}
template <vstart_t v_start, typename vertex_t>
constexpr bool is_vertex_valid(vertex_t vid) {
    return vid != invalid_vertex<v_start, vertex_t>();
}

}  // namespace utils

template <typename T>
auto print_helper(const T& v) {
    if constexpr (std::is_class_v<T> && utils::HasToStrMethod<T>::value) return v.ToString();
    else return v;
}

template <typename... Ts>
void printx(Ts&&... args) {
    (std::cout << ... << print_helper(args)) << std::endl;
}
template <typename... Ts>
void LOG_INFO(Ts&&... args) {
    std::cout << "[INFO]  ";
    printx(std::forward<Ts>(args)...);
}
template <typename... Ts>
void LOG_WARNING(Ts&&... args) {
    std::cout << "[WARN]  ";
    printx(std::forward<Ts>(args)...);
}
template <typename... Ts>
void LOG_DEBUG(Ts&&... args) {
#ifdef DEBUG_LOG
    std::cout << "[DEBUG] ";
    printx(std::forward<Ts>(args)...);
#else
    ((void)args, ...);
#endif
}
template <typename... Ts>
[[noreturn]] void LOG_ERROR(Ts&&... args) {
    std::cout << "[ERROR] ";
    printx(std::forward<Ts>(args)...);
    std::exit(1);
}

}  // namespace graph_one

// ---------------------------------------------------------------------------------------------------
// Command-line parsing: the subset of CLI11's interface the graph_one apps use, written here from its
// documented behaviour (options "-x,--long", required options, boolean flags, --help).
namespace CLI {

class Option {
public:
    Option(std::string names, std::string desc, bool flag) : desc_(std::move(desc)), flag_(flag) {
        size_t p = 0;
        while (p <= names.size()) {
            const size_t q = names.find(',', p);
            std::string nm = names.substr(p, q == std::string::npos ? std::string::npos : q - p);
            if (!nm.empty()) names_.push_back(nm);
            if (q == std::string::npos) break;
            p = q + 1;
        }
    }
    virtual ~Option() = default;
    Option* required(bool r = true) {
        required_ = r;
        return this;
    }
    bool matches(const std::string& a) const {
        for (const auto& n : names_)
            if (n == a) return true;
        return false;
    }
    virtual bool assign(const std::string& value) = 0;
    const std::vector<std::string>& names() const { return names_; }
    bool is_flag() const { return flag_; }
    bool is_required() const { return required_; }
    bool seen = false;
    const std::string& desc() const { return desc_; }

private:
    std::vector<std::string> names_;
    std::string desc_;
    bool flag_;
    bool required_ = false;
};

template <typename T>
class TypedOption : public Option {
public:
    TypedOption(std::string names, T& target, std::string desc, bool flag)
        : Option(std::move(names), std::move(desc), flag), target_(target) {}
    bool assign(const std::string& value) override {
        if constexpr (std::is_same_v<T, bool>) {
            target_ = !(value == "false" || value == "0");
            return true;
        } else if constexpr (std::is_same_v<T, std::string>) {
            target_ = value;
            return true;
        } else {
            std::istringstream is(value);
            is >> target_;
            return !is.fail();
        }
    }

private:
    T& target_;
};

struct ParseError {
    int code;
    std::string msg;
};

class App {
public:
    explicit App(std::string desc = "") : desc_(std::move(desc)) {}
    template <typename T>
    Option* add_option(const std::string& names, T& target, const std::string& desc = "") {
        opts_.emplace_back(new TypedOption<T>(names, target, desc, false));
        return opts_.back().get();
    }
    Option* add_flag(const std::string& names, bool& target, const std::string& desc = "") {
        opts_.emplace_back(new TypedOption<bool>(names, target, desc, true));
        return opts_.back().get();
    }
    void parse(int argc, char** argv) {
        for (int i = 1; i < argc; ++i) {
            std::string a = argv[i];
            if (a == "-h" || a == "--help") throw ParseError{0, help()};
            std::string value;
            bool has_value = false;
            const size_t eq = a.find('=');
            if (a.rfind("--", 0) == 0 && eq != std::string::npos) {
                value = a.substr(eq + 1);
                a = a.substr(0, eq);
                has_value = true;
            }
            Option* o = find(a);
            if (!o) throw ParseError{109, "The following argument was not expected: " + a};
            if (o->is_flag()) {
                if (!o->assign(has_value ? value : "true")) throw ParseError{105, "bad value for " + a};
            } else {
                if (!has_value) {
                    if (i + 1 >= argc) throw ParseError{107, a + " requires an argument"};
                    value = argv[++i];
                }
                if (!o->assign(value)) throw ParseError{105, "could not convert '" + value + "' for " + a};
            }
            o->seen = true;
        }
        for (const auto& o : opts_)
            if (o->is_required() && !o->seen) throw ParseError{106, o->names().back() + " is required"};
    }
    int exit(const ParseError& e) const {
        if (e.code == 0) {
            std::cout << e.msg;
            return 0;
        }
        std::cerr << e.msg << "\nRun with --help for more information." << std::endl;
        return e.code;
    }
    std::string help() const {
        std::string h = desc_.empty() ? "" : desc_ + "\n";
        h += "Options:\n";
        for (const auto& o : opts_) {
            std::string nm;
            for (const auto& n : o->names()) nm += (nm.empty() ? "" : ",") + n;
            h += "  " + nm + (o->is_required() ? " REQUIRED" : "") + "  " + o->desc() + "\n";
        }
        return h;
    }

private:
    Option* find(const std::string& a) {
        for (auto& o : opts_)
            if (o->matches(a)) return o.get();
        return nullptr;
    }
    std::string desc_;
    std::vector<std::unique_ptr<Option>> opts_;
};

}  // namespace CLI

#define CLI11_PARSE(app, argc, argv)                                                                   \
    try {                                                                                              \
        (app).parse((argc), (argv));                                                                   \
    } catch (const CLI::ParseError& e) {                                                               \
        return (app).exit(e);                                                                          \
    }
