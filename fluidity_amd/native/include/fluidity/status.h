// Status values of the HIP operator layer.
//
// The reference operators return absl::Status / absl::StatusOr with fixed code meanings
// (native/include/fluidity/host/transfers_3d.h:17-22, mac_pressure_3d.h:95-100). Abseil is not part
// of this build, so the layer carries its own small Status type whose codes are numerically equal to
// absl::StatusCode; an embedding that has Abseil converts with absl::Status(absl::StatusCode(int(code())), "").
#pragma once

#include <optional>
#include <utility>

namespace fluidity {

enum class StatusCode : int {
  kOk = 0,
  kInvalidArgument = 3,
  kResourceExhausted = 8,
  kFailedPrecondition = 9,
  kAborted = 10,
  kOutOfRange = 11,
  kUnimplemented = 12,
  kInternal = 13,
};

// Code-only status, like the reference's code-only errors.
class Status {
 public:
  Status() = default;
  explicit Status(StatusCode code) : code_(code) {}
  bool ok() const { return code_ == StatusCode::kOk; }
  StatusCode code() const { return code_; }
  int raw_code() const { return static_cast<int>(code_); }
  friend bool operator==(const Status& a, const Status& b) { return a.code_ == b.code_; }

 private:
  StatusCode code_ = StatusCode::kOk;
};

inline Status OkStatus() { return Status(); }
// Status of a C-ABI return value (the values are the StatusCode numbers).
inline Status FromCode(int code) { return Status(static_cast<StatusCode>(code)); }

// A value or a non-OK status.
template <class T>
class StatusOr {
 public:
  StatusOr(Status status) : status_(status) {}  // NOLINT: implicit like absl::StatusOr
  StatusOr(T value) : value_(std::move(value)) {}  // NOLINT
  bool ok() const { return status_.ok(); }
  const Status& status() const { return status_; }
  T& value() & { return *value_; }
  const T& value() const& { return *value_; }
  T&& value() && { return std::move(*value_); }
  T& operator*() & { return *value_; }
  const T& operator*() const& { return *value_; }
  T* operator->() { return &*value_; }
  const T* operator->() const { return &*value_; }

 private:
  Status status_;
  std::optional<T> value_;
};

}  // namespace fluidity
