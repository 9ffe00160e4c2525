#include "transfer.h"

#include "../../include/sundog_api.h"

#include <cmath>
#include <cstdio>
#include <stdexcept>

namespace sdk {

namespace {

// SMPTE ST 2084 constants, exact rationals as published.
constexpr double kM1 = 2610.0 / 16384.0;
constexpr double kM2 = 2523.0 / 4096.0 * 128.0;
constexpr double kC1 = 3424.0 / 4096.0;
constexpr double kC2 = 2413.0 / 4096.0 * 32.0;
constexpr double kC3 = 2392.0 / 4096.0 * 32.0;

float clamp01(float v) { return std::fmin(std::fmax(v, 0.0f), 1.0f); }  // NaN -> 0

// RGB -> XYZ matrix of a colour space from its primaries' and white's xy
// chromaticities (the standard derivation: columns are the primaries' XYZ
// scaled so that RGB = 1 maps to the white point with Y = 1).
void rgbToXyz(const double xy[4][2], double m[3][3]) {
  double P[3][3];
  for (int i = 0; i < 3; i++) {
    double x = xy[i][0], y = xy[i][1];
    P[0][i] = x / y;
    P[1][i] = 1.0;
    P[2][i] = (1.0 - x - y) / y;
  }
  double W[3] = {xy[3][0] / xy[3][1], 1.0, (1.0 - xy[3][0] - xy[3][1]) / xy[3][1]};
  // solve P * S = W (Cramer's rule)
  auto det3 = [](const double a[3][3]) {
    return a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
           a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
  };
  double d = det3(P), S[3];
  for (int c = 0; c < 3; c++) {
    double Q[3][3];
    for (int r = 0; r < 3; r++)
      for (int k = 0; k < 3; k++) Q[r][k] = k == c ? W[r] : P[r][k];
    S[c] = det3(Q) / d;
  }
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) m[r][c] = P[r][c] * S[c];
}

void invert3(const double a[3][3], double o[3][3]) {
  double d = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
             a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) {
      int r1 = (c + 1) % 3, r2 = (c + 2) % 3, c1 = (r + 1) % 3, c2 = (r + 2) % 3;
      o[r][c] = (a[r1][c1] * a[r2][c2] - a[r1][c2] * a[r2][c1]) / d;
    }
}

struct Gamut {
  float m[3][3];
  Gamut() {
    // BT.709 and BT.2020 primaries (R, G, B) and the D65 white point
    const double p709[4][2] = {{0.640, 0.330}, {0.300, 0.600}, {0.150, 0.060}, {0.3127, 0.3290}};
    const double p2020[4][2] = {{0.708, 0.292}, {0.170, 0.797}, {0.131, 0.046}, {0.3127, 0.3290}};
    double a[3][3], b[3][3], bi[3][3];
    rgbToXyz(p709, a);
    rgbToXyz(p2020, b);
    invert3(b, bi);
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) {
        double s = 0.0;
        for (int k = 0; k < 3; k++) s += bi[r][k] * a[k][c];
        m[r][c] = (float)s;
      }
  }
};

const Gamut& gamut() {
  static const Gamut g;
  return g;
}

}  // namespace

float pqOetf(float linear) {
  double y = clamp01(linear);
  double p = std::pow(y, kM1);
  return (float)std::pow((kC1 + kC2 * p) / (1.0 + kC3 * p), kM2);
}

float pqEotf(float code) {
  double e = clamp01(code);
  double p = std::pow(e, 1.0 / kM2);
  double num = std::fmax(p - kC1, 0.0), den = kC2 - kC3 * p;
  return (float)std::pow(num / den, 1.0 / kM1);
}

void bt709To2020(const float in[3], float out[3]) {
  const Gamut& g = gamut();
  for (int r = 0; r < 3; r++) out[r] = g.m[r][0] * in[0] + g.m[r][1] * in[1] + g.m[r][2] * in[2];
}

void encodePq12(const float* rgba, int w, int h, float exposure, uint16_t* out) {
  const float scale = std::exp2(exposure);
  const float toPq = kPqReferenceWhiteNits / kPqPeakNits;  // linear 1.0 -> 203 nits
  for (size_t i = 0; i < (size_t)w * h; i++) {
    float c[3] = {rgba[4 * i] * scale, rgba[4 * i + 1] * scale, rgba[4 * i + 2] * scale}, o[3];
    bt709To2020(c, o);
    for (int k = 0; k < 3; k++) out[3 * i + k] = (uint16_t)std::lround(pqOetf(o[k] * toPq) * 4095.0f);
  }
}

void writePpm12(const std::string& path, const uint16_t* rgb, int w, int h) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  fprintf(f, "P6\n%d %d\n4095\n", w, h);
  std::vector<unsigned char> row((size_t)w * 6);
  for (int y = 0; y < h; y++) {
    for (size_t k = 0; k < (size_t)w * 3; k++) {
      uint16_t v = rgb[(size_t)y * w * 3 + k];
      row[2 * k] = (unsigned char)(v >> 8);
      row[2 * k + 1] = (unsigned char)(v & 0xFF);
    }
    fwrite(row.data(), 1, row.size(), f);
  }
  if (fclose(f) != 0) throw std::runtime_error("cannot write " + path);
}

}  // namespace sdk

extern "C" float sundog_pq_oetf(float linear) { return sdk::pqOetf(linear); }
extern "C" float sundog_pq_eotf(float code) { return sdk::pqEotf(code); }
extern "C" void sundog_bt709_to_2020(const float in[3], float out[3]) { sdk::bt709To2020(in, out); }
extern "C" int sundog_encode_pq12(const float* rgba, int width, int height, float exposure, uint16_t* out) {
  if (!rgba || !out || width <= 0 || height <= 0) return SUNDOG_ERROR;
  sdk::encodePq12(rgba, width, height, exposure, out);
  return SUNDOG_OK;
}
