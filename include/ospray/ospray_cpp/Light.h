// OSPRay C++ wrapper: Light (ospray_cpp/Light.h of OSPRay 3.x).
#pragma once

#include "ManagedObject.h"

namespace ospray {
namespace cpp {

class Light : public ManagedObject<OSPLight, OSP_LIGHT>
{
 public:
  Light(const std::string &type) : ManagedObject<OSPLight, OSP_LIGHT>(ospNewLight(type.c_str())) {}
  Light(OSPLight existing = nullptr) : ManagedObject<OSPLight, OSP_LIGHT>(existing) {}
};

static_assert(sizeof(Light) == sizeof(OSPLight), "cpp::Light holds only its handle");

} // namespace cpp

OSPTYPEFOR_SPECIALIZATION(cpp::Light, OSP_LIGHT);

} // namespace ospray
