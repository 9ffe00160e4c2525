// Host asset readers (own implementations; no third-party parsers).
#include "assets.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

namespace sdk {

std::string resolvePath(const std::string& baseDir, const std::string& p) {
  if (!p.empty() && p[0] != '/') return baseDir + "/" + p;
  return p;
}

// ---------------------------------------------------------------------------
// OBJ
// ---------------------------------------------------------------------------

namespace {

struct Corner {
  uint32_t v;
  int32_t vt;
};

// Parses "v", "v/vt", "v//vn", "v/vt/vn" with 1-based or negative indices.
bool parseCorner(const char* s, size_t nv, size_t nvt, Corner& c) {
  char* end = nullptr;
  long vi = strtol(s, &end, 10);
  if (end == s) return false;
  long vt = 0;
  if (*end == '/') {
    const char* q = end + 1;
    if (*q != '/') {
      char* e2 = nullptr;
      vt = strtol(q, &e2, 10);
      if (e2 == q) vt = 0;
    }
  }
  long v0 = vi > 0 ? vi - 1 : (long)nv + vi;
  long t0 = vt > 0 ? vt - 1 : (vt < 0 ? (long)nvt + vt : -1);
  if (v0 < 0) return false;
  c.v = (uint32_t)v0;
  c.vt = (int32_t)t0;
  return true;
}

}  // namespace

HostMesh loadObj(const std::string& path, bool smoothNormals, const std::string& usemtl) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("OBJ load failed: " + path + ": cannot open file");
  std::vector<vec3> pos;
  std::vector<vec2> tex;
  std::vector<Corner> corners;
  std::string curMtl;
  bool groupSeen = usemtl.empty();
  std::string line;
  std::vector<Corner> poly;
  while (std::getline(in, line)) {
    const char* s = line.c_str();
    while (*s == ' ' || *s == '\t') s++;
    if (s[0] == 'v' && s[1] == ' ') {
      vec3 p{0, 0, 0};
      if (sscanf(s + 2, "%f %f %f", &p.x, &p.y, &p.z) != 3)
        throw std::runtime_error("OBJ load failed: " + path + ": bad vertex line");
      pos.push_back(p);
    } else if (s[0] == 'v' && s[1] == 't' && (s[2] == ' ' || s[2] == '\t')) {
      vec2 t{0, 0};
      sscanf(s + 3, "%f %f", &t.x, &t.y);
      tex.push_back(t);
    } else if (strncmp(s, "usemtl", 6) == 0 && (s[6] == ' ' || s[6] == '\t')) {
      std::istringstream ss(s + 7);
      ss >> curMtl;
      if (!usemtl.empty() && curMtl == usemtl) groupSeen = true;
    } else if (s[0] == 'f' && (s[1] == ' ' || s[1] == '\t')) {
      poly.clear();
      std::istringstream ss(s + 2);
      std::string tok;
      while (ss >> tok) {
        Corner c;
        if (!parseCorner(tok.c_str(), pos.size(), tex.size(), c))
          throw std::runtime_error("OBJ load failed: " + path + ": bad face token");
        poly.push_back(c);
      }
      if (!usemtl.empty() && curMtl != usemtl) continue;
      for (size_t k = 1; k + 1 < poly.size(); k++) {  // fan triangulation
        corners.push_back(poly[0]);
        corners.push_back(poly[k]);
        corners.push_back(poly[k + 1]);
      }
    }
  }
  if (!groupSeen) throw std::runtime_error("OBJ usemtl group not found: " + usemtl + " in " + path);
  bool allUv = !tex.empty();
  for (const Corner& c : corners) {
    if (c.v >= pos.size()) throw std::runtime_error("OBJ has out-of-range vertex index: " + path);
    if (c.vt >= 0 && (size_t)c.vt >= tex.size())
      throw std::runtime_error("OBJ has out-of-range texcoord index: " + path);
    if (c.vt < 0) allUv = false;
  }
  if (corners.empty())
    throw std::runtime_error(usemtl.empty() ? "OBJ has no triangles: " + path
                                            : "OBJ usemtl group has no triangles: " + usemtl +
                                                  " in " + path);

  HostMesh m;
  std::vector<uint32_t> posOf;
  m.indices.resize(corners.size());
  if (allUv) {
    std::unordered_map<uint64_t, uint32_t> remap;
    remap.reserve(pos.size() * 2);
    for (size_t c = 0; c < corners.size(); c++) {
      uint64_t key = ((uint64_t)corners[c].v << 32) | (uint32_t)corners[c].vt;
      auto it = remap.find(key);
      uint32_t id;
      if (it == remap.end()) {
        id = (uint32_t)m.positions.size();
        remap.emplace(key, id);
        m.positions.push_back(pos[corners[c].v]);
        m.uvs.push_back(tex[corners[c].vt]);
        posOf.push_back(corners[c].v);
      } else {
        id = it->second;
      }
      m.indices[c] = id;
    }
  } else {
    m.positions = pos;
    posOf.resize(pos.size());
    for (size_t i = 0; i < pos.size(); i++) posOf[i] = (uint32_t)i;
    for (size_t c = 0; c < corners.size(); c++) m.indices[c] = corners[c].v;
  }
  if (smoothNormals) {
    std::vector<vec3> acc(pos.size(), v3(0.0f));
    for (size_t t = 0; t < m.numTris(); t++) {
      uint32_t a = m.indices[3 * t], b = m.indices[3 * t + 1], c = m.indices[3 * t + 2];
      vec3 n = cross(m.positions[b] - m.positions[a], m.positions[c] - m.positions[a]);
      acc[posOf[a]] += n;
      acc[posOf[b]] += n;
      acc[posOf[c]] += n;
    }
    m.normals.resize(m.positions.size());
    for (size_t i = 0; i < m.positions.size(); i++) {
      vec3 n = acc[posOf[i]];
      float l = length(n);
      m.normals[i] = l > 1e-12f ? n / l : v3(0.0f, 1.0f, 0.0f);
    }
  }
  return m;
}

// ---------------------------------------------------------------------------
// Radiance RGBE
// ---------------------------------------------------------------------------

static void rgbeToFloat(const uint8_t* e, float* out) {
  if (e[3] == 0) {
    out[0] = out[1] = out[2] = 0.0f;
  } else {
    float f = ldexpf(1.0f, (int)e[3] - (128 + 8));
    out[0] = (float)e[0] * f;
    out[1] = (float)e[1] * f;
    out[2] = (float)e[2] * f;
  }
  out[3] = 1.0f;
}

HostImage loadRadianceHdr(const std::string& path) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("cannot load envmap '" + path + "'");
  auto fail = [&](const char* why) {
    fclose(f);
    throw std::runtime_error("cannot load envmap '" + path + "': " + why);
  };
  char line[512];
  if (!fgets(line, sizeof line, f)) fail("empty file");
  if (strncmp(line, "#?", 2) != 0) fail("not a Radiance file");
  bool rgbe = false;
  while (fgets(line, sizeof line, f)) {
    if (line[0] == '\n' || line[0] == '\r') break;
    if (strncmp(line, "FORMAT=32-bit_rle_rgbe", 22) == 0) rgbe = true;
    if (strncmp(line, "FORMAT=32-bit_rle_xyze", 22) == 0) fail("XYZE is not supported");
  }
  (void)rgbe;
  int w = 0, h = 0;
  if (!fgets(line, sizeof line, f) || sscanf(line, "-Y %d +X %d", &h, &w) != 2 || w <= 0 || h <= 0)
    fail("unsupported resolution line (need -Y H +X W)");
  HostImage img;
  img.width = w;
  img.height = h;
  img.rgba.resize((size_t)w * h * 4);
  std::vector<uint8_t> scan((size_t)w * 4);
  for (int y = 0; y < h; y++) {
    uint8_t hdr[4];
    if (fread(hdr, 1, 4, f) != 4) fail("truncated scanline");
    bool rle = w >= 8 && w < 32768 && hdr[0] == 2 && hdr[1] == 2 && !(hdr[2] & 0x80) &&
               ((hdr[2] << 8) | hdr[3]) == w;
    if (!rle) {
      memcpy(scan.data(), hdr, 4);
      if (w > 1 && fread(scan.data() + 4, 1, (size_t)(w - 1) * 4, f) != (size_t)(w - 1) * 4)
        fail("truncated flat scanline");
    } else {
      for (int c = 0; c < 4; c++) {
        int x = 0;
        while (x < w) {
          int n = fgetc(f);
          if (n == EOF) fail("truncated RLE scanline");
          if (n > 128) {
            n -= 128;
            int v = fgetc(f);
            if (v == EOF || x + n > w) fail("bad RLE run");
            for (int k = 0; k < n; k++) scan[(size_t)(x++) * 4 + c] = (uint8_t)v;
          } else {
            if (n == 0 || x + n > w) fail("bad RLE literal");
            for (int k = 0; k < n; k++) {
              int v = fgetc(f);
              if (v == EOF) fail("truncated RLE literal");
              scan[(size_t)(x++) * 4 + c] = (uint8_t)v;
            }
          }
        }
      }
    }
    for (int x = 0; x < w; x++) rgbeToFloat(&scan[(size_t)x * 4], &img.rgba[((size_t)y * w + x) * 4]);
  }
  fclose(f);
  return img;
}

// ---------------------------------------------------------------------------
// Image textures (PFM / P6 PPM)
// ---------------------------------------------------------------------------

static float srgbToLinear(float c) {
  return c <= 0.04045f ? c / 12.92f : powf((c + 0.055f) / 1.055f, 2.4f);
}

HostImage loadTextureImage(const std::string& path, bool srgb) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("cannot load texture '" + path + "'");
  char magic[3] = {0, 0, 0};
  if (fscanf(f, "%2s", magic) != 1) {
    fclose(f);
    throw std::runtime_error("texture '" + path + "': unreadable header");
  }
  HostImage img;
  int w = 0, h = 0;
  if (strcmp(magic, "PF") == 0 || strcmp(magic, "Pf") == 0) {
    float scale = 0;
    if (fscanf(f, "%d %d %f", &w, &h, &scale) != 3 || w <= 0 || h <= 0) {
      fclose(f);
      throw std::runtime_error("texture '" + path + "': bad PFM header");
    }
    fgetc(f);
    int ch = magic[1] == 'F' ? 3 : 1;
    std::vector<float> raw((size_t)w * h * ch);
    if (fread(raw.data(), 4, raw.size(), f) != raw.size()) {
      fclose(f);
      throw std::runtime_error("texture '" + path + "': truncated PFM");
    }
    if (scale > 0.0f) {  // a positive scale marks big-endian samples
      for (float& v : raw) {
        uint32_t u;
        memcpy(&u, &v, 4);
        u = __builtin_bswap32(u);
        memcpy(&v, &u, 4);
      }
    }
    img.width = w;
    img.height = h;
    img.rgba.resize((size_t)w * h * 4);
    for (int y = 0; y < h; y++)  // PFM stores bottom row first
      for (int x = 0; x < w; x++) {
        const float* s = &raw[((size_t)(h - 1 - y) * w + x) * ch];
        float* d = &img.rgba[((size_t)y * w + x) * 4];
        d[0] = s[0];
        d[1] = s[ch > 1 ? 1 : 0];
        d[2] = s[ch > 1 ? 2 : 0];
        d[3] = 1.0f;
      }
  } else if (strcmp(magic, "P6") == 0) {
    int maxv = 0;
    if (fscanf(f, "%d %d %d", &w, &h, &maxv) != 3 || w <= 0 || h <= 0 || maxv != 255) {
      fclose(f);
      throw std::runtime_error("texture '" + path + "': bad PPM header");
    }
    fgetc(f);
    std::vector<uint8_t> raw((size_t)w * h * 3);
    if (fread(raw.data(), 1, raw.size(), f) != raw.size()) {
      fclose(f);
      throw std::runtime_error("texture '" + path + "': truncated PPM");
    }
    img.width = w;
    img.height = h;
    img.rgba.resize((size_t)w * h * 4);
    for (size_t i = 0; i < (size_t)w * h; i++) {
      for (int c = 0; c < 3; c++) {
        float v = raw[i * 3 + c] / 255.0f;
        img.rgba[i * 4 + c] = srgb ? srgbToLinear(v) : v;
      }
      img.rgba[i * 4 + 3] = 1.0f;
    }
  } else {
    fclose(f);
    throw std::runtime_error("texture '" + path +
                             "': unsupported format (this build reads .pfm and P6 .ppm; "
                             "AVIF decoding is not built)");
  }
  fclose(f);
  return img;
}

}  // namespace sdk
