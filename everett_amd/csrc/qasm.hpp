// OpenQASM 3.0 text <-> Circuit: the reference's minimal subset (src/qasm/).
// The reader is the serving wire format: a request arrives as program text.
#pragma once

#include <cstddef>
#include <string>

#include "circuit.hpp"

namespace ev {

// src/qasm/parse.rs:811-819 (Circuit::from_qasm, src/circuit.rs:324-326).
// Throws Error{EV_E_QASM, a = line, b = col} with the reference's message;
// the text is "OpenQASM 3 error at <line>:<col>: <message>" (src/error.rs:82-84).
Circuit parse_qasm(const char *src, size_t len);

// src/qasm/emit.rs:17-37 (Circuit::to_qasm, src/circuit.rs:315-317).  Throws
// Error{EV_E_QASM, 0, 0} for a gate with no OpenQASM 3 name or a nested
// IfClassic.
std::string emit_qasm(const Circuit &c);

}  // namespace ev
