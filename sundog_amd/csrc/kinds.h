// Enumerations shared by the host scene model and the kernels; the values
// are the reference's (device/params.h:12-18) and cross the ABI unchanged.
#pragma once

#include <stdint.h>

namespace sdk {

enum GeomKind : int { GK_SPHERE = 0, GK_RECT, GK_DISK, GK_CYLINDER, GK_PARABOLA, GK_MESH };
enum TexKind : int { TX_SOLID = 0, TX_IMAGE, TX_CHECKER, TX_GRID };
enum MatKind : int { MT_LAMBERT = 0, MT_METAL, MT_DIELECTRIC, MT_EMISSIVE, MT_WATER, MT_PLASTIC };
enum LightKind : int { LT_RECT = 0, LT_DISK, LT_SPHERE, LT_POINT, LT_DISTANT, LT_MESH };
enum BgKind : int { BG_SOLID = 0, BG_GRADIENT, BG_ENVMAP };

constexpr uint16_t kMatNone = 0xFFFF;  // pass-through face

}  // namespace sdk
