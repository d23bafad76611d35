// OSPRay C++ wrapper: Group (ospray_cpp/Group.h of OSPRay 3.x).
#pragma once

#include "ManagedObject.h"

namespace ospray {
namespace cpp {

class Group : public ManagedObject<OSPGroup, OSP_GROUP>
{
 public:
  Group() : ManagedObject<OSPGroup, OSP_GROUP>(ospNewGroup()) {}
  Group(OSPGroup existing) : ManagedObject<OSPGroup, OSP_GROUP>(existing) {}
};

static_assert(sizeof(Group) == sizeof(OSPGroup), "cpp::Group holds only its handle");

} // namespace cpp

OSPTYPEFOR_SPECIALIZATION(cpp::Group, OSP_GROUP);

} // namespace ospray
