// OSPRay C++ wrapper: Future (ospray_cpp/Future.h of OSPRay 3.x).
#pragma once

#include "ManagedObject.h"

namespace ospray {
namespace cpp {

class Future : public ManagedObject<OSPFuture, OSP_FUTURE>
{
 public:
  Future(OSPFuture existing = nullptr) : ManagedObject<OSPFuture, OSP_FUTURE>(existing) {}
  bool isReady(OSPSyncEvent e = OSP_TASK_FINISHED) { return ospIsReady(handle(), e) != 0; }
  void wait(OSPSyncEvent e = OSP_TASK_FINISHED) { ospWait(handle(), e); }
  void cancel() { ospCancel(handle()); }
  float progress() { return ospGetProgress(handle()); }
  float duration() { return ospGetTaskDuration(handle()); }
};

static_assert(sizeof(Future) == sizeof(OSPFuture), "cpp::Future holds only its handle");

} // namespace cpp

OSPTYPEFOR_SPECIALIZATION(cpp::Future, OSP_FUTURE);

} // namespace ospray
