// OSPRay C++ wrapper: Geometry (ospray_cpp/Geometry.h of OSPRay 3.x).
#pragma once

#include "ManagedObject.h"

namespace ospray {
namespace cpp {

class Geometry : public ManagedObject<OSPGeometry, OSP_GEOMETRY>
{
 public:
  Geometry(const std::string &type) : ManagedObject<OSPGeometry, OSP_GEOMETRY>(ospNewGeometry(type.c_str())) {}
  Geometry(OSPGeometry existing = nullptr) : ManagedObject<OSPGeometry, OSP_GEOMETRY>(existing) {}
};

static_assert(sizeof(Geometry) == sizeof(OSPGeometry), "cpp::Geometry holds only its handle");

} // namespace cpp

OSPTYPEFOR_SPECIALIZATION(cpp::Geometry, OSP_GEOMETRY);

} // namespace ospray
