// Op-counting scalar for the CPU oracle (test infrastructure only).
//
// Replays the oracle's arithmetic on a value type that tallies every
// floating-point operation it performs, so that the algorithmic work per
// (K, q) pair can be counted branch-accurately on real inputs. Values carry a
// `var` flag: literals, the element (A, I, Z) and the projectile mass are
// launch invariants, K and q are per-pair variables. An operation is tallied
// only when one of its operands depends on K or q; work on invariants alone
// is per-element set-up, done once per launch, and is not per-pair work.
// Transcendental calls are tallied by kind; their cost in FP64 VALU
// operations is applied afterwards from a separate cost model.
#pragma once

#include <cmath>
#include <cstdint>

namespace noa_oracle {

enum OpKind {
    OP_ADD = 0,  // + and -
    OP_MUL,
    OP_DIV,
    OP_SQRT,
    OP_LOG,
    OP_EXP,
    OP_LOG10,
    OP_POW,
    OP_CMP,
    OP_NKINDS
};

struct OpTally {
    uint64_t n[OP_NKINDS];
};

extern thread_local OpTally g_tally;

struct Counted {
    double v;
    bool var;
    Counted() : v(0.0), var(false) {}
    Counted(double x) : v(x), var(false) {}  // NOLINT: implicit from literals, like double
    Counted(double x, bool is_var) : v(x), var(is_var) {}
    Counted& operator+=(const Counted& o);
    Counted& operator/=(const Counted& o);
};

inline Counted tally2(OpKind k, double r, const Counted& a, const Counted& b) {
    const bool var = a.var || b.var;
    if (var) g_tally.n[k]++;
    return Counted(r, var);
}
inline Counted tally1(OpKind k, double r, const Counted& a) {
    if (a.var) g_tally.n[k]++;
    return Counted(r, a.var);
}

inline Counted operator+(Counted a, Counted b) { return tally2(OP_ADD, a.v + b.v, a, b); }
inline Counted operator-(Counted a, Counted b) { return tally2(OP_ADD, a.v - b.v, a, b); }
inline Counted operator*(Counted a, Counted b) { return tally2(OP_MUL, a.v * b.v, a, b); }
inline Counted operator/(Counted a, Counted b) { return tally2(OP_DIV, a.v / b.v, a, b); }
inline Counted operator-(Counted a) { return Counted(-a.v, a.var); }  // a source modifier on VALU
inline Counted& Counted::operator+=(const Counted& o) { return *this = *this + o; }
inline Counted& Counted::operator/=(const Counted& o) { return *this = *this / o; }
inline bool cmp_tally(const Counted& a, const Counted& b) {
    if (a.var || b.var) g_tally.n[OP_CMP]++;
    return true;
}
inline bool operator<(Counted a, Counted b) { return cmp_tally(a, b) && a.v < b.v; }
inline bool operator<=(Counted a, Counted b) { return cmp_tally(a, b) && a.v <= b.v; }
inline bool operator>(Counted a, Counted b) { return cmp_tally(a, b) && a.v > b.v; }
inline bool operator>=(Counted a, Counted b) { return cmp_tally(a, b) && a.v >= b.v; }
inline Counted log(Counted a) { return tally1(OP_LOG, std::log(a.v), a); }
inline Counted exp(Counted a) { return tally1(OP_EXP, std::exp(a.v), a); }
inline Counted log10(Counted a) { return tally1(OP_LOG10, std::log10(a.v), a); }
inline Counted sqrt(Counted a) { return tally1(OP_SQRT, std::sqrt(a.v), a); }
inline Counted pow(Counted a, Counted b) { return tally2(OP_POW, std::pow(a.v, b.v), a, b); }

}  // namespace noa_oracle
