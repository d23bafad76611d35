// OSPRay C++ wrapper: Texture (ospray_cpp/Texture.h of OSPRay 3.x).
#pragma once

#include "ManagedObject.h"

namespace ospray {
namespace cpp {

class Texture : public ManagedObject<OSPTexture, OSP_TEXTURE>
{
 public:
  Texture(const std::string &type) : ManagedObject<OSPTexture, OSP_TEXTURE>(ospNewTexture(type.c_str())) {}
  Texture(OSPTexture existing = nullptr) : ManagedObject<OSPTexture, OSP_TEXTURE>(existing) {}
};

static_assert(sizeof(Texture) == sizeof(OSPTexture), "cpp::Texture holds only its handle");

} // namespace cpp

OSPTYPEFOR_SPECIALIZATION(cpp::Texture, OSP_TEXTURE);

} // namespace ospray
