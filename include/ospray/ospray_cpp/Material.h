// OSPRay C++ wrapper: Material (ospray_cpp/Material.h of OSPRay 3.x).
#pragma once

#include "ManagedObject.h"

namespace ospray {
namespace cpp {

class Material : public ManagedObject<OSPMaterial, OSP_MATERIAL>
{
 public:
  Material(const std::string &type) : ManagedObject<OSPMaterial, OSP_MATERIAL>(ospNewMaterial(type.c_str())) {}
  Material(OSPMaterial existing = nullptr) : ManagedObject<OSPMaterial, OSP_MATERIAL>(existing) {}
};

static_assert(sizeof(Material) == sizeof(OSPMaterial), "cpp::Material holds only its handle");

} // namespace cpp

OSPTYPEFOR_SPECIALIZATION(cpp::Material, OSP_MATERIAL);

} // namespace ospray
