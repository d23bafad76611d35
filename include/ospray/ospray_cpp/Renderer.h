// OSPRay C++ wrapper: Renderer (ospray_cpp/Renderer.h of OSPRay 3.x).
#pragma once

#include "ManagedObject.h"

namespace ospray {
namespace cpp {

class Renderer : public ManagedObject<OSPRenderer, OSP_RENDERER>
{
 public:
  Renderer(const std::string &type) : ManagedObject<OSPRenderer, OSP_RENDERER>(ospNewRenderer(type.c_str())) {}
  Renderer(OSPRenderer existing = nullptr) : ManagedObject<OSPRenderer, OSP_RENDERER>(existing) {}
};

static_assert(sizeof(Renderer) == sizeof(OSPRenderer), "cpp::Renderer holds only its handle");

} // namespace cpp

OSPTYPEFOR_SPECIALIZATION(cpp::Renderer, OSP_RENDERER);

} // namespace ospray
