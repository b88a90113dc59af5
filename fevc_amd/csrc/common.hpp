// Shared host-side types for the MI355X leave-out engine.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace vk {

// Status codes of the reference ABI (rust/crates/vckss-core/src/error.rs:8-47).
enum Code : int32_t {
  kOk = 0,
  kUserBreak = 1,
  kAbiMismatch = 10,
  kUnsupported = 11,
  kInvalidInput = 20,
  kInvalidIdentifier = 21,
  kInvalidWeight = 22,
  kInvalidTargetWeight = 23,
  kGraphEmpty = 30,
  kGraphUnidentified = 31,
  kGraphCertificate = 32,
  kResourceLimit = 40,
  kAllocationFailed = 41,
  kCmgSetup = 50,
  kCmgApply = 51,
  kPcgCurvature = 60,
  kPcgPreconditioner = 61,
  kPcgStagnation = 62,
  kPcgMaxIterations = 63,
  kFullResidual = 64,
  kRngContract = 70,
  kCorrectionNonFinite = 71,
  kAccounting = 72,
  kJlaMoment = 73,
  kJlaConstraint = 74,
  kNonestimable = 75,
  kBlockInverse = 76,
  kTargetCentering = 77,
  kTargetIdentity = 78,
  kNonfiniteCorrected = 79,
  kStaleContext = 80,
  kContextPoisoned = 81,
  kSingularInformation = 82,
  kInverseResidual = 83,
  kInvariant = 90,
  kPanic = 99,
};

struct Error : std::runtime_error {
  int32_t code;
  Error(int32_t c, const std::string &phase, const std::string &msg)
      : std::runtime_error(phase + ": " + msg), code(c) {}
};

[[noreturn]] inline void fail(int32_t code, const std::string &phase, const std::string &msg) {
  throw Error(code, phase, msg);
}

// Cooperative interruption of a call (ffi_engine.rs:1412-1503).  The caller's
// poll runs synchronously on the calling thread at named checkpoints, once
// every `interval` checkpoints (the first checkpoint polls): 0 continues, 1 is
// a user break (status 1), any other value an invariant failure (90).  A null
// poll needs a null context and a zero interval; a poll needs a positive one.
struct Interrupt {
  int32_t (*poll)(void *) = nullptr;
  void *context = nullptr;
  uint32_t interval = 0, remaining = 0;
  uint64_t polls = 0;

  Interrupt() = default;
  Interrupt(int32_t (*p)(void *), void *c, uint32_t every) : poll(p), context(c), interval(every) {
    if (!poll && (context || interval))
      fail(kInvalidInput, "engine_interrupt", "a null interrupt callback requires null context and zero interval");
    if (poll && !interval)
      fail(kInvalidInput, "engine_interrupt", "a nonnull interrupt callback requires a positive checkpoint interval");
  }
  void checkpoint(const char *phase) {
    if (!poll) return;
    if (remaining) {
      --remaining;
      return;
    }
    remaining = interval - 1;
    ++polls;
    const int32_t v = poll(context);
    if (v == 0) return;
    if (v == 1) fail(kUserBreak, phase, "user requested interruption");
    fail(kInvariant, "engine_interrupt", "interrupt callback returned unsupported status " + std::to_string(v));
  }
};

// Neumaier compensated accumulator (engine.rs:327-352, graph.rs:863-871).
struct Neumaier {
  double sum = 0.0, corr = 0.0;
  inline void add(double v) {
    double next = sum + v;
    corr += (std::fabs(sum) >= std::fabs(v)) ? ((sum - next) + v) : ((v - next) + sum);
    sum = next;
  }
  inline double value() const { return sum + corr; }
};

// Kahan accumulator returning the running sum (problem.rs:1250-1267).
struct Kahan {
  double sum = 0.0, corr = 0.0;
  inline void add(double v) {
    double adjusted = v - corr;
    double next = sum + adjusted;
    corr = (next - sum) - adjusted;
    sum = next;
  }
};

inline int64_t total_key(double x) {
  int64_t b;
  std::memcpy(&b, &x, 8);
  return b ^ static_cast<int64_t>(static_cast<uint64_t>(b >> 63) >> 1);
}
inline double canonical_zero(double x) { return x == 0.0 ? 0.0 : x; }
inline uint64_t bits_of(double x) {
  uint64_t b;
  std::memcpy(&b, &x, 8);
  return b;
}

}  // namespace vk
