// Circuit text front end: the line format read by the reference's
// CliffordCircuit (plaquette_simulator/src/CliffordCircuit.hpp:32-66),
// parsed straight into an op list for programs.
//
// A line is an operation name followed by qubit indices, separated by
// whitespace.  Names: H S X Y Z (one op per listed qubit, in order), CX CZ
// (one op per consecutive (control, target) pair), M (one measurement per
// listed qubit, each into a new record, bias 1/2 -- MeasureQubits,
// CliffordState.hpp:543-549).  Indices are read as unsigned integers until
// the first token that is not one; the rest of the line is ignored, as the
// reference's `line_stream >> qubit` loop does.  Any other name, including
// the empty name of a blank line, is an error with the reference's message
// "Unsupported operation: <name>".
#include <stdint.h>

#include <cctype>
#include <cstring>
#include <string>
#include <vector>

#include "plaquette_amd.h"
#include "pq_capi_internal.h"

namespace {

enum : uint32_t { kNotQubit = 0xFFFFFFFFu };

// Reads one whitespace-separated token starting at *pos within the line
// [*pos, end).  Returns false at the end of the line.
bool next_token(const char *text, uint64_t end, uint64_t *pos, std::string *tok) {
  uint64_t i = *pos;
  while (i < end && std::isspace((unsigned char)text[i])) ++i;
  if (i >= end) {
    *pos = i;
    return false;
  }
  const uint64_t s = i;
  while (i < end && !std::isspace((unsigned char)text[i])) ++i;
  tok->assign(text + s, i - s);
  *pos = i;
  return true;
}

// A qubit token: decimal digits, optionally signed.  A leading '-' gives an
// index no state has (the reference's unsigned read wraps it), so it fails
// the range check when the ops are validated.  Returns false for anything
// that does not start with a number; *whole is false when characters follow
// the number ("3,4" reads 3 and ends the line's indices, as a stream read
// stops at the ',').
bool parse_qubit(const std::string &tok, uint32_t *q, bool *whole) {
  size_t i = 0;
  bool neg = false;
  if (i < tok.size() && (tok[i] == '+' || tok[i] == '-')) neg = tok[i++] == '-';
  if (i >= tok.size() || !std::isdigit((unsigned char)tok[i])) return false;
  uint64_t v = 0;
  for (; i < tok.size() && std::isdigit((unsigned char)tok[i]); ++i) {
    v = v * 10 + (uint64_t)(tok[i] - '0');
    if (v > 0xFFFFFFFEull) v = 0xFFFFFFFEull;  // saturate: out of range anyway
  }
  *q = (neg && v) ? kNotQubit - 1 : (uint32_t)v;
  *whole = i == tok.size();
  return true;
}

}  // namespace

extern "C" int pq_circuit_parse(const char *text, uint64_t len, uint32_t first_record,
                                pq_op *ops, uint64_t capacity, uint64_t *nops,
                                uint64_t *nlines, uint64_t *qubits_used) {
  if (!text && len) return pq_set_error(PQ_ERR_ARG, "null circuit text");
  uint64_t count = 0, lines = 0, qmax = 0;
  uint32_t rec = first_record;
  const uint32_t half = pq_coin_threshold(0.5f);
  auto emit = [&](uint32_t code, uint32_t q0, uint32_t q1, uint32_t slot) {
    if (ops && count < capacity) {
      pq_op &o = ops[count];
      std::memset(&o, 0, sizeof(o));
      o.code = code;
      o.q0 = q0;
      o.q1 = q1;
      o.slot = slot;
    }
    ++count;
    if ((uint64_t)q0 + 1 > qmax) qmax = (uint64_t)q0 + 1;
    if ((code == PQ_OP_CX || code == PQ_OP_CZ) && (uint64_t)q1 + 1 > qmax) qmax = (uint64_t)q1 + 1;
  };
  uint64_t pos = 0;
  while (pos < len) {
    // one line: [pos, eol); std::getline semantics (a final line without
    // a newline counts, a trailing newline does not open an empty line)
    uint64_t eol = pos;
    while (eol < len && text[eol] != '\n') ++eol;
    const uint64_t line_no = lines + 1;
    std::string name, tok;
    uint64_t p = pos;
    next_token(text, eol, &p, &name);
    uint32_t code = 0;
    if (name == "H") code = PQ_OP_H;
    else if (name == "S") code = PQ_OP_S;
    else if (name == "X") code = PQ_OP_X;
    else if (name == "Y") code = PQ_OP_Y;
    else if (name == "Z") code = PQ_OP_Z;
    else if (name == "CX") code = PQ_OP_CX;
    else if (name == "CZ") code = PQ_OP_CZ;
    else if (name == "M") code = PQ_OP_M;
    else
      return pq_set_error(PQ_ERR_ARG, ("Unsupported operation: " + name).c_str());
    std::vector<uint32_t> qs;
    uint32_t q = 0;
    bool whole = true;
    while (whole && next_token(text, eol, &p, &tok) && parse_qubit(tok, &q, &whole))
      qs.push_back(q);
    if ((code == PQ_OP_CX || code == PQ_OP_CZ) && (qs.size() % 2))
      return pq_set_error(PQ_ERR_ARG, ("line " + std::to_string(line_no) + ": " + name +
                                       " needs (control, target) pairs").c_str());
    if (code == PQ_OP_CX || code == PQ_OP_CZ) {
      for (size_t i = 0; i < qs.size(); i += 2) emit(code, qs[i], qs[i + 1], 0);
    } else if (code == PQ_OP_M) {
      for (uint32_t t : qs) {
        emit(code, t, 0, rec++);
        if (ops && count <= capacity) ops[count - 1].t0 = half;
      }
    } else {
      for (uint32_t t : qs) emit(code, t, 0, 0);
    }
    ++lines;
    pos = eol + 1;
  }
  if (nops) *nops = count;
  if (nlines) *nlines = lines;
  if (qubits_used) *qubits_used = qmax;
  if (ops && count > capacity) return pq_set_error(PQ_ERR_ARG, "op buffer too small");
  return PQ_OK;
}
