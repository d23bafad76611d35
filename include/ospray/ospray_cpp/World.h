// OSPRay C++ wrapper: World (ospray_cpp/World.h of OSPRay 3.x).
#pragma once

#include "Instance.h"
#include "Light.h"

namespace ospray {
namespace cpp {

class World : public ManagedObject<OSPWorld, OSP_WORLD>
{
 public:
  World() : ManagedObject<OSPWorld, OSP_WORLD>(ospNewWorld()) {}
  World(OSPWorld existing) : ManagedObject<OSPWorld, OSP_WORLD>(existing) {}
};

static_assert(sizeof(World) == sizeof(OSPWorld), "cpp::World holds only its handle");

} // namespace cpp

OSPTYPEFOR_SPECIALIZATION(cpp::World, OSP_WORLD);

} // namespace ospray
