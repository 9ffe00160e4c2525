// HDR output transfer of the film: exposure -> BT.709 to BT.2020 gamut ->
// 203-nit reference white -> SMPTE ST 2084 (PQ) -> 12-bit code values, the
// chain of the reference's writer (src/film.cpp:96-115, src/transfer.h).
// GPU-free.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace sdk {

constexpr float kPqReferenceWhiteNits = 203.0f;  // BT.2408 graphics white
constexpr float kPqPeakNits = 10000.0f;

// PQ inverse EOTF: linear (1.0 = 10,000 nits) -> non-linear [0, 1]. Inputs
// are clamped to [0, 1] first, so negative values and NaN encode as 0 and
// +inf as 1.
float pqOetf(float linear);
// PQ EOTF: non-linear [0, 1] (clamped) -> linear.
float pqEotf(float code);
// Linear BT.709 RGB -> linear BT.2020 RGB (D65 white preserved).
void bt709To2020(const float in[3], float out[3]);
// w x h RGBA float film -> w x h x 3 12-bit PQ / BT.2020 code values.
void encodePq12(const float* rgba, int w, int h, float exposure, uint16_t* out);
// Binary PPM (P6) with maxval 4095: two big-endian bytes per sample.
void writePpm12(const std::string& path, const uint16_t* rgb, int w, int h);

}  // namespace sdk
