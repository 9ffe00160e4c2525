// Render preparation shared by the GPU session and the flat scene export:
// render overrides, asset loading, mesh-light patching and per-object
// inverse transforms. GPU-free.
#pragma once

#include <vector>

#include "../../include/sundog_api.h"
#include "assets.h"
#include "scene.h"

namespace sdk {

struct MeshLightPatch {
  int lightId = -1;
  int objectId = -1;
  std::vector<vec3> worldPos;
  std::vector<float> cdf;  // numTris + 1
  float area = 0.0f;
  float ngSign = 1.0f;
};

struct Prepared {
  Scene scene;                        // after overrides
  std::vector<HostMesh> meshes;
  std::vector<HostImage> texImages;   // index = texture id (empty for non-image)
  HostImage env;                      // empty unless envmap background
  std::vector<Affine> inverses;       // per object, world -> object
  std::vector<MeshLightPatch> meshLights;
  double loadMs = 0.0;
};

// Applies render-option overrides in place (reference src/capi_render.cpp:73-83).
void applyOverrides(Scene& s, const sundog_render_options* opt);

// Loads assets and derives render-time data from an already-overridden scene.
void prepareScene(const Scene& s, Prepared& out);

// world -> object inverse of a row-major 3x4 affine, computed in double and
// rounded once to float.
Affine invertAffine(const Affine& m);

void setLastError(const std::string& msg);

}  // namespace sdk
