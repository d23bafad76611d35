// OSPRay C++ wrapper: Volume (ospray_cpp/Volume.h of OSPRay 3.x).
#pragma once

#include "ManagedObject.h"

namespace ospray {
namespace cpp {

class Volume : public ManagedObject<OSPVolume, OSP_VOLUME>
{
 public:
  Volume(const std::string &type) : ManagedObject<OSPVolume, OSP_VOLUME>(ospNewVolume(type.c_str())) {}
  Volume(OSPVolume existing = nullptr) : ManagedObject<OSPVolume, OSP_VOLUME>(existing) {}
};

static_assert(sizeof(Volume) == sizeof(OSPVolume), "cpp::Volume holds only its handle");

} // namespace cpp

OSPTYPEFOR_SPECIALIZATION(cpp::Volume, OSP_VOLUME);

} // namespace ospray
