/* Minimal C application of the OSPRay C API: builds a small scene, renders
 * it with the "pathtracer" renderer and writes the accumulated image.  It
 * uses nothing but include/ospray/ospray.h and ospray_util.h, the way an
 * application written for OSPRay 3.x does (compare the reference's
 * apps/ospTutorial/ospTutorial.c), so it checks that the drop-in boundary
 * compiles and runs unchanged.
 *
 * usage: pathtracer_app [--osp:load-modules=hip --osp:device=hip] out.ppm frames
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "ospray/ospray.h"
#include "ospray/ospray_util.h"

static void onError(void *user, OSPError e, const char *msg)
{
  (void)user;
  fprintf(stderr, "OSPRay error %d: %s\n", (int)e, msg);
  exit(1);
}

int main(int argc, const char **argv)
{
  if (ospInit(&argc, argv) != OSP_NO_ERROR)
    return 1;
  OSPDevice dev = ospGetCurrentDevice();
  ospDeviceSetErrorCallback(dev, onError, NULL);
  ospDeviceRelease(dev);
  const char *out = argc > 1 ? argv[1] : "out.ppm";
  const int frames = argc > 2 ? atoi(argv[2]) : 4;
  const int W = 160, H = 120;

  /* two triangles and a ground quad */
  float vertex[] = {-1.f, -1.f, 3.f, -1.f, 1.f, 3.f, 1.f, -1.f, 3.f, 0.1f, 0.1f, 0.3f, -3.f, -1.5f, -1.f, 3.f,
      -1.5f, -1.f, 3.f, -1.5f, 6.f, -3.f, -1.5f, 6.f};
  unsigned int tris[] = {0, 1, 2, 1, 2, 3};
  unsigned int quad[] = {4, 5, 6, 7};
  OSPGeometry mesh = ospNewGeometry("mesh");
  OSPData v = ospNewSharedData1D(vertex, OSP_VEC3F, 8);
  OSPData it = ospNewSharedData1D(tris, OSP_VEC3UI, 2);
  ospSetObject(mesh, "vertex.position", v);
  ospSetObject(mesh, "index", it);
  ospCommit(mesh);
  OSPGeometry ground = ospNewGeometry("mesh");
  OSPData iq = ospNewSharedData1D(quad, OSP_VEC4UI, 1);
  ospSetObject(ground, "vertex.position", v);
  ospSetObject(ground, "index", iq);
  ospCommit(ground);
  ospRelease(v);
  ospRelease(it);
  ospRelease(iq);

  OSPMaterial red = ospNewMaterial("obj");
  ospSetVec3f(red, "kd", 0.8f, 0.2f, 0.1f);
  ospCommit(red);
  OSPMaterial metal = ospNewMaterial("metal");
  ospSetFloat(metal, "roughness", 0.2f);
  ospCommit(metal);

  OSPGeometricModel m0 = ospNewGeometricModel(mesh);
  ospSetObject(m0, "material", red);
  ospCommit(m0);
  OSPGeometricModel m1 = ospNewGeometricModel(ground);
  ospSetObject(m1, "material", metal);
  ospCommit(m1);
  OSPGroup group = ospNewGroup();
  OSPGeometricModel models[2] = {m0, m1};
  OSPData md = ospNewSharedData1D(models, OSP_GEOMETRIC_MODEL, 2);
  OSPData mdc = ospNewData1D(OSP_GEOMETRIC_MODEL, 2);
  ospCopyData1D(md, mdc, 0);
  ospSetObject(group, "geometry", mdc);
  ospCommit(group);
  ospRelease(md);
  ospRelease(mdc);
  OSPInstance inst = ospNewInstance(group);
  ospCommit(inst);

  OSPLight lights[2];
  lights[0] = ospNewLight("quad");
  ospSetVec3f(lights[0], "position", -1.f, 3.f, 1.f);
  ospSetVec3f(lights[0], "edge1", 0.f, 0.f, 2.f);
  ospSetVec3f(lights[0], "edge2", 2.f, 0.f, 0.f);
  ospSetFloat(lights[0], "intensity", 5.f);
  ospCommit(lights[0]);
  lights[1] = ospNewLight("ambient");
  ospSetFloat(lights[1], "intensity", 0.3f);
  ospCommit(lights[1]);

  OSPWorld world = ospNewWorld();
  ospSetObjectAsData(world, "instance", OSP_INSTANCE, inst);
  OSPData ld = ospNewSharedData1D(lights, OSP_LIGHT, 2);
  ospSetObject(world, "light", ld);
  ospCommit(world);
  ospRelease(ld);

  OSPCamera camera = ospNewCamera("perspective");
  ospSetFloat(camera, "aspect", (float)W / (float)H);
  ospSetVec3f(camera, "position", 0.f, 0.f, -2.f);
  ospSetVec3f(camera, "direction", 0.1f, 0.f, 1.f);
  ospSetVec3f(camera, "up", 0.f, 1.f, 0.f);
  ospCommit(camera);

  OSPRenderer renderer = ospNewRenderer("pathtracer");
  ospSetInt(renderer, "pixelSamples", 2);
  ospSetInt(renderer, "maxPathLength", 6);
  ospCommit(renderer);

  OSPFrameBuffer fb = ospNewFrameBuffer(W, H, OSP_FB_SRGBA, OSP_FB_COLOR | OSP_FB_ACCUM | OSP_FB_VARIANCE);
  ospResetAccumulation(fb);
  float variance = 0.f;
  for (int f = 0; f < frames; f++)
    variance = ospRenderFrameBlocking(fb, renderer, camera, world);

  const unsigned char *px = (const unsigned char *)ospMapFrameBuffer(fb, OSP_FB_COLOR);
  FILE *fp = fopen(out, "wb");
  if (!fp || !px)
    return 2;
  fprintf(fp, "P6\n%d %d\n255\n", W, H);
  unsigned long long sum = 0;
  for (int y = H - 1; y >= 0; y--)
    for (int x = 0; x < W; x++) {
      fwrite(px + 4 * (y * W + x), 1, 3, fp);
      sum += px[4 * (y * W + x)] + px[4 * (y * W + x) + 1] + px[4 * (y * W + x) + 2];
    }
  fclose(fp);
  ospUnmapFrameBuffer(px, fb);
  printf("frames %d variance %g checksum %llu\n", frames, variance, sum);

  ospRelease(fb);
  ospRelease(renderer);
  ospRelease(camera);
  ospRelease(world);
  ospRelease(lights[0]);
  ospRelease(lights[1]);
  ospRelease(inst);
  ospRelease(group);
  ospRelease(m0);
  ospRelease(m1);
  ospRelease(red);
  ospRelease(metal);
  ospRelease(mesh);
  ospRelease(ground);
  ospShutdown();
  return 0;
}
