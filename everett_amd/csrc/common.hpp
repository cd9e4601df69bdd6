// Shared host-side types of the everett MI355X engine: complex scalars, the
// error model behind the C ABI's status codes, and HIP call checking.
#pragma once

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>

#include "../../include/everett.h"

namespace ev {

// src/complex.rs:22-29 -- the reference's Complex64 (two f64).
struct c128 {
    double re, im;
};
static_assert(sizeof(c128) == sizeof(ev_c128), "c128 must match ev_c128");

inline c128 cadd(c128 a, c128 b) { return {a.re + b.re, a.im + b.im}; }
// src/complex.rs:115-125: (fma(a.re,b.re,-(a.im*b.im)), fma(a.re,b.im,a.im*b.re)).
inline c128 cmul(c128 a, c128 b)
{
    return {__builtin_fma(a.re, b.re, -(a.im * b.im)), __builtin_fma(a.re, b.im, a.im * b.re)};
}
inline c128 cconj(c128 a) { return {a.re, -a.im}; }
inline bool ceq(c128 a, c128 b) { return a.re == b.re && a.im == b.im; }

// An error carried to the C ABI boundary, where it becomes a status code.
struct Error : std::runtime_error {
    int code;
    uint64_t a, b;
    Error(int code_, std::string msg, uint64_t a_ = 0, uint64_t b_ = 0)
        : std::runtime_error(std::move(msg)), code(code_), a(a_), b(b_) {}
};

[[noreturn]] void throw_hip(int err, const char *what, const char *file, int line);

#define EV_HIP(call)                                                        \
    do {                                                                    \
        int ev_e_ = static_cast<int>(call);                                 \
        if (ev_e_ != 0) ::ev::throw_hip(ev_e_, #call, __FILE__, __LINE__);  \
    } while (0)

#define EV_REQUIRE(cond, code, msg)                                         \
    do {                                                                    \
        if (!(cond)) throw ::ev::Error((code), (msg));                      \
    } while (0)

// Largest state this build addresses: 2^34 amplitudes (256 GiB as c128 fits
// the 288 GB of one MI355X).  Indices are 64-bit throughout.
constexpr uint64_t kMaxQubits = 34;

}  // namespace ev
