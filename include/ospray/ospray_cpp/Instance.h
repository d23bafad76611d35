// OSPRay C++ wrapper: Instance (ospray_cpp/Instance.h of OSPRay 3.x).
#pragma once

#include "Group.h"

namespace ospray {
namespace cpp {

class Instance : public ManagedObject<OSPInstance, OSP_INSTANCE>
{
 public:
  Instance(const Group &group) : Instance(group.handle()) {}
  Instance(OSPGroup group) : ManagedObject<OSPInstance, OSP_INSTANCE>(ospNewInstance(group)) {}
  Instance(OSPInstance existing = nullptr) : ManagedObject<OSPInstance, OSP_INSTANCE>(existing) {}
};

static_assert(sizeof(Instance) == sizeof(OSPInstance), "cpp::Instance holds only its handle");

} // namespace cpp

OSPTYPEFOR_SPECIALIZATION(cpp::Instance, OSP_INSTANCE);

} // namespace ospray
