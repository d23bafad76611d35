// The C application's scene (pathtracer_app.c) written against the OSPRay
// C++ wrapper (ospray/ospray_cpp.h), the way an application written for
// OSPRay 3.x's ospray::cpp classes does (compare the reference's
// apps/ospTutorial/ospTutorial.cpp).  It renders the same frames and
// prints the same checksum line, then a pick and the world's bounds.
//
// usage: pathtracer_cpp_app [--osp:load-modules=hip --osp:device=hip] out.ppm frames
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ospray/ospray_cpp.h"

namespace app {
// the application's own math types, registered with OSPTypeFor below
struct vec3f
{
  float x, y, z;
};
struct vec3ui
{
  unsigned int x, y, z;
};
struct vec4ui
{
  unsigned int x, y, z, w;
};
struct box3f
{
  vec3f lower, upper;
};
} // namespace app

namespace ospray {
OSPTYPEFOR_SPECIALIZATION(app::vec3f, OSP_VEC3F);
OSPTYPEFOR_SPECIALIZATION(app::vec3ui, OSP_VEC3UI);
OSPTYPEFOR_SPECIALIZATION(app::vec4ui, OSP_VEC4UI);
OSPTYPEFOR_SPECIALIZATION(app::box3f, OSP_BOX3F);
OSPTYPEFOR_DEFINITION(app::vec3f);
} // namespace ospray

using namespace ospray;
using app::vec3f;

int main(int argc, const char **argv)
{
  if (ospInit(&argc, argv) != OSP_NO_ERROR)
    return 1;
  int rc = 0;
  {
    cpp::Device dev = cpp::Device::current();
    dev.setErrorCallback([](void *, OSPError e, const char *msg) {
      std::fprintf(stderr, "OSPRay error %d: %s\n", (int)e, msg);
      std::exit(1);
    });
    const char *out = argc > 1 ? argv[1] : "out.ppm";
    const int frames = argc > 2 ? std::atoi(argv[2]) : 4;
    const int W = 160, H = 120;

    // two triangles and a ground quad
    std::vector<vec3f> vertex = {{-1.f, -1.f, 3.f}, {-1.f, 1.f, 3.f}, {1.f, -1.f, 3.f}, {0.1f, 0.1f, 0.3f},
        {-3.f, -1.5f, -1.f}, {3.f, -1.5f, -1.f}, {3.f, -1.5f, 6.f}, {-3.f, -1.5f, 6.f}};
    std::vector<app::vec3ui> tris = {{0, 1, 2}, {1, 2, 3}};
    std::vector<app::vec4ui> quad = {{4, 5, 6, 7}};

    cpp::Geometry mesh("mesh");
    mesh.setParam("vertex.position", cpp::SharedData(vertex));
    mesh.setParam("index", cpp::CopiedData(tris));
    mesh.commit();
    cpp::Geometry ground("mesh");
    ground.setParam("vertex.position", cpp::SharedData(vertex));
    ground.setParam("index", cpp::MovedData(quad)); // quad is empty from here on
    ground.commit();

    cpp::Material red("obj");
    red.setParam("kd", vec3f{0.8f, 0.2f, 0.1f});
    red.commit();
    cpp::Material metal("metal");
    metal.setParam("roughness", 0.2f);
    metal.commit();

    cpp::GeometricModel m0(mesh);
    m0.setParam("material", red);
    m0.commit();
    cpp::GeometricModel m1(ground);
    m1.setParam("material", metal);
    m1.commit();

    cpp::Group group;
    group.setParam("geometry", cpp::CopiedData(std::vector<cpp::GeometricModel>{m0, m1}));
    group.commit();
    cpp::Instance instance(group);
    instance.commit();

    cpp::Light quadLight("quad");
    quadLight.setParam("position", vec3f{-1.f, 3.f, 1.f});
    quadLight.setParam("edge1", vec3f{0.f, 0.f, 2.f});
    quadLight.setParam("edge2", vec3f{2.f, 0.f, 0.f});
    quadLight.setParam("intensity", 5.f);
    quadLight.commit();
    cpp::Light ambient("ambient");
    ambient.setParam("intensity", 0.3f);
    ambient.commit();
    std::vector<cpp::Light> lights = {quadLight, ambient};

    cpp::World world;
    world.setParam("instance", cpp::CopiedData(instance));
    world.setParam("light", cpp::SharedData(lights));
    world.commit();

    cpp::Camera camera("perspective");
    camera.setParam("aspect", (float)W / (float)H);
    camera.setParam("position", vec3f{0.f, 0.f, -2.f});
    camera.setParam("direction", vec3f{0.1f, 0.f, 1.f});
    camera.setParam("up", vec3f{0.f, 1.f, 0.f});
    camera.commit();

    cpp::Renderer renderer("pathtracer");
    renderer.setParam("pixelSamples", 2);
    renderer.setParam("maxPathLength", 6);
    renderer.commit();

    cpp::FrameBuffer fb(W, H, OSP_FB_SRGBA, OSP_FB_COLOR | OSP_FB_ACCUM | OSP_FB_VARIANCE);
    fb.clear();
    float variance = 0.f;
    for (int f = 0; f < frames; f++) {
      cpp::Future future = fb.renderFrame(renderer, camera, world);
      future.wait();
      if (!future.isReady())
        return 3;
      variance = fb.variance();
    }

    const unsigned char *px = (const unsigned char *)fb.map(OSP_FB_COLOR);
    std::FILE *fp = std::fopen(out, "wb");
    if (!fp || !px)
      return 2;
    std::fprintf(fp, "P6\n%d %d\n255\n", W, H);
    unsigned long long sum = 0;
    for (int y = H - 1; y >= 0; y--)
      for (int x = 0; x < W; x++) {
        std::fwrite(px + 4 * (y * W + x), 1, 3, fp);
        sum += px[4 * (y * W + x)] + px[4 * (y * W + x) + 1] + px[4 * (y * W + x) + 2];
      }
    std::fclose(fp);
    fb.unmap((void *)px);
    std::printf("frames %d variance %g checksum %llu\n", frames, variance, sum);

    // screen point (0.6, 0.4) sees the first red triangle at distance 5.04
    cpp::PickResult p = fb.pick(renderer, camera, world, 0.6f, 0.4f);
    std::printf("pick hit %d model %d primID %u\n", (int)p.hasHit, p.hasHit && p.model == m0 ? 0 : -1, p.primID);
    const app::box3f b = world.getBounds<app::box3f>();
    std::printf("bounds %g %g %g %g %g %g\n", b.lower.x, b.lower.y, b.lower.z, b.upper.x, b.upper.y, b.upper.z);
    rc = p.hasHit ? 0 : 4;
  } // every wrapper releases its object here, before shutdown
  ospShutdown();
  return rc;
}
