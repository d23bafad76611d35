// OSPRay C++ wrapper: FrameBuffer and PickResult (ospray_cpp/FrameBuffer.h
// of OSPRay 3.x).
#pragma once

#include "Camera.h"
#include "Future.h"
#include "GeometricModel.h"
#include "Instance.h"
#include "Renderer.h"
#include "World.h"

namespace ospray {
namespace cpp {

struct PickResult
{
  bool hasHit{false};
  float worldPosition[3];
  Instance instance{(OSPInstance) nullptr};
  GeometricModel model{(OSPGeometricModel) nullptr};
  uint32_t primID{0};
};

class FrameBuffer : public ManagedObject<OSPFrameBuffer, OSP_FRAMEBUFFER>
{
 public:
  FrameBuffer() = default; // no device object
  FrameBuffer(int size_x, int size_y, OSPFrameBufferFormat format = OSP_FB_SRGBA, int channels = OSP_FB_COLOR)
      : ManagedObject<OSPFrameBuffer, OSP_FRAMEBUFFER>(ospNewFrameBuffer(size_x, size_y, format, channels))
  {
  }
  FrameBuffer(OSPFrameBuffer existing) : ManagedObject<OSPFrameBuffer, OSP_FRAMEBUFFER>(existing) {}

  void resetAccumulation() const { ospResetAccumulation(handle()); }
  Future renderFrame(const Renderer &renderer, const Camera &camera, const World &world) const
  {
    return Future(ospRenderFrame(handle(), renderer.handle(), camera.handle(), world.handle()));
  }
  PickResult pick(const Renderer &renderer, const Camera &camera, const World &world, float screenPos_x,
      float screenPos_y) const
  {
    OSPPickResult p;
    ospPick(&p, handle(), renderer.handle(), camera.handle(), world.handle(), screenPos_x, screenPos_y);
    PickResult res;
    res.hasHit = p.hasHit != 0;
    if (res.hasHit) {
      for (int k = 0; k < 3; k++)
        res.worldPosition[k] = p.worldPosition[k];
      res.instance = Instance(p.instance); // takes over the references ospPick returned
      res.model = GeometricModel(p.model);
      res.primID = p.primID;
    }
    return res;
  }
  void *map(OSPFrameBufferChannel channel) const { return const_cast<void *>(ospMapFrameBuffer(handle(), channel)); }
  void unmap(void *ptr) const { ospUnmapFrameBuffer(ptr, handle()); }
  void clear() const { ospResetAccumulation(handle()); }
  float variance() const { return ospGetVariance(handle()); }
};

static_assert(sizeof(FrameBuffer) == sizeof(OSPFrameBuffer), "cpp::FrameBuffer holds only its handle");

} // namespace cpp

OSPTYPEFOR_SPECIALIZATION(cpp::FrameBuffer, OSP_FRAMEBUFFER);

} // namespace ospray
