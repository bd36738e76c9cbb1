// Execution-space tags (reference: hydra/device/System.h, hydra/host/System.h).
// hydra::device is the MI355X; hydra::host is host memory (data only: every
// phase-space computation runs on the device).
#pragma once

namespace hydra {
namespace device {
struct sys_t {};
constexpr sys_t sys{};
}  // namespace device
namespace host {
struct sys_t {};
constexpr sys_t sys{};
}  // namespace host
}  // namespace hydra
