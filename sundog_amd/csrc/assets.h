// Host asset readers: Wavefront OBJ meshes, Radiance RGBE (.hdr)
// environment images and PFM / binary PPM image textures. GPU-free.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "sd_math.h"

namespace sdk {

struct HostMesh {
  std::vector<vec3> positions;
  std::vector<uint32_t> indices;  // 3 per triangle
  std::vector<vec3> normals;      // empty when smooth normals are off
  std::vector<vec2> uvs;          // empty when the OBJ has no complete vt set
  size_t numTris() const { return indices.size() / 3; }
};

// Semantics of the reference loader (src/mesh_obj.cpp:13-145): polygons fan-
// triangulated, optional usemtl group filter, (v, vt) seam re-indexing when
// every corner has a texcoord, OBJ normals ignored, area-weighted smooth
// normals aggregated per original position.
HostMesh loadObj(const std::string& path, bool smoothNormals, const std::string& usemtl);

// Linear float RGBA, row 0 = top of the file.
struct HostImage {
  int width = 0, height = 0;
  std::vector<float> rgba;
};

HostImage loadRadianceHdr(const std::string& path);
// Image textures: .pfm (linear) or .ppm P6 (8-bit, sRGB-decoded when srgb).
HostImage loadTextureImage(const std::string& path, bool srgb);

std::string resolvePath(const std::string& baseDir, const std::string& p);

}  // namespace sdk
