// flop_count.hh -- operation-counting scalar for the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY.  Force-included (-include) ahead of
// oracle_capi.cpp to build liboracle_count.so: every `double` of the oracle
// (and of the ABI header's structs) becomes `cd`, a standard-layout wrapper
// of one double (same size, alignment and calling convention, so the C API
// and the caller's arrays are unchanged) whose arithmetic operators count
// the operations they perform:
//   add  additions and subtractions     mul  multiplications
//   div  divisions                      sqrt square roots
//   tr   pow / log / exp                (fabs, fmin, fmax, copysign and
//                                        comparisons are not counted)
// Counts are kept per pass (ORC_PASS / ORC_SUB / ORC_METRIC in the oracle's
// substep passes) and the library runs single-threaded, so plain counters
// suffice.  No fused multiply-add exists in the oracle (it is built with
// -ffp-contract=off), so add + mul is the algorithm's flop count.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>
#include <omp.h>

namespace orcflop {
enum { F_ADD = 0, F_MUL, F_DIV, F_SQRT, F_TR, F_N };
constexpr int NPASS = 20;
inline int64_t g_counts[NPASS][F_N] = {};
inline int g_pass = 0;
inline void tick(int k) { ++g_counts[g_pass][k]; }
// Scoped re-tagging (ORC_SUB, ORC_METRIC): restores the enclosing pass.
struct Sub {
    int saved;
    explicit Sub(int id) : saved(g_pass) { g_pass = id; }
    ~Sub() { g_pass = saved; }
};
inline int metric_id() { return g_pass < 10 ? 10 + g_pass : g_pass; }
}  // namespace orcflop

struct cd {
    double v;
    cd() = default;
    constexpr cd(double x) : v(x) {}
    explicit constexpr operator double() const { return v; }
    explicit constexpr operator float() const { return (float)v; }
    explicit constexpr operator int() const { return (int)v; }
    explicit constexpr operator long() const { return (long)v; }
    explicit constexpr operator long long() const { return (long long)v; }
    explicit constexpr operator unsigned long() const { return (unsigned long)v; }
    explicit constexpr operator bool() const { return v != 0.0; }
    constexpr cd operator-() const { return cd(-v); }
    constexpr cd operator+() const { return *this; }
    cd& operator+=(cd o) { orcflop::tick(orcflop::F_ADD); v += o.v; return *this; }
    cd& operator-=(cd o) { orcflop::tick(orcflop::F_ADD); v -= o.v; return *this; }
    cd& operator*=(cd o) { orcflop::tick(orcflop::F_MUL); v *= o.v; return *this; }
    cd& operator/=(cd o) { orcflop::tick(orcflop::F_DIV); v /= o.v; return *this; }
};
static_assert(sizeof(cd) == sizeof(double) && alignof(cd) == alignof(double), "cd must alias double");

inline cd operator+(cd a, cd b) { orcflop::tick(orcflop::F_ADD); return cd(a.v + b.v); }
inline cd operator-(cd a, cd b) { orcflop::tick(orcflop::F_ADD); return cd(a.v - b.v); }
inline cd operator*(cd a, cd b) { orcflop::tick(orcflop::F_MUL); return cd(a.v * b.v); }
inline cd operator/(cd a, cd b) { orcflop::tick(orcflop::F_DIV); return cd(a.v / b.v); }
inline bool operator<(cd a, cd b) { return a.v < b.v; }
inline bool operator>(cd a, cd b) { return a.v > b.v; }
inline bool operator<=(cd a, cd b) { return a.v <= b.v; }
inline bool operator>=(cd a, cd b) { return a.v >= b.v; }
inline bool operator==(cd a, cd b) { return a.v == b.v; }
inline bool operator!=(cd a, cd b) { return a.v != b.v; }

namespace std {
inline cd sqrt(cd a) { orcflop::tick(orcflop::F_SQRT); return cd(::sqrt(a.v)); }
inline cd fabs(cd a) { return cd(::fabs(a.v)); }
inline cd abs(cd a) { return cd(::fabs(a.v)); }
inline cd copysign(cd a, cd b) { return cd(::copysign(a.v, b.v)); }
inline cd fmax(cd a, cd b) { return cd(::fmax(a.v, b.v)); }
inline cd fmin(cd a, cd b) { return cd(::fmin(a.v, b.v)); }
inline cd pow(cd a, cd b) { orcflop::tick(orcflop::F_TR); return cd(::pow(a.v, b.v)); }
inline cd log(cd a) { orcflop::tick(orcflop::F_TR); return cd(::log(a.v)); }
inline cd exp(cd a) { orcflop::tick(orcflop::F_TR); return cd(::exp(a.v)); }
inline cd floor(cd a) { return cd(::floor(a.v)); }
inline cd min(double a, cd b) { return b.v < a ? b : cd(a); }
inline cd min(cd a, double b) { return b < a.v ? cd(b) : a; }
inline cd max(double a, cd b) { return a < b.v ? b : cd(a); }
inline cd max(cd a, double b) { return a.v < b ? cd(b) : a; }
inline bool isfinite(cd a) { return std::isfinite(a.v); }
inline bool isnan(cd a) { return std::isnan(a.v); }
template <>
struct numeric_limits<cd> {
    static constexpr bool is_specialized = true;
    static constexpr cd epsilon() { return cd(numeric_limits<double>::epsilon()); }
    static constexpr cd max() { return cd(numeric_limits<double>::max()); }
    static constexpr cd min() { return cd(numeric_limits<double>::min()); }
    static constexpr cd lowest() { return cd(numeric_limits<double>::lowest()); }
    static constexpr cd infinity() { return cd(numeric_limits<double>::infinity()); }
    static constexpr cd quiet_NaN() { return cd(numeric_limits<double>::quiet_NaN()); }
};
}  // namespace std
using std::sqrt;
using std::fabs;
using std::copysign;
using std::fmax;
using std::fmin;
using std::pow;
using std::log;
using std::exp;

#define ORC_PASS(id) (orcflop::g_pass = (id))
#define ORC_SUB(id) orcflop::Sub orc_sub_guard_(id)
#define ORC_METRIC() orcflop::Sub orc_sub_guard_(orcflop::metric_id())
#define double cd
