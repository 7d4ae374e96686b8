// Spectral helpers of the lens camera (generate.metal:122-157): the CIE 1931
// 2-degree colour matching functions at 5 nm from 380 to 780 nm (CIE 15:2004
// tables, 4 decimals) and the spectral weight of a wavelength drawn
// uniformly over that range.
#pragma once
#include "common.h"

namespace rm {

// CIE XYZ to linear sRGB primaries, D65 (IEC 61966-2-1)
RM_HD vec3 xyz_to_rgb(vec3 v) {
    return vec3((3.240479f * v.x + -1.537150f * v.y) + -0.498535f * v.z,
                (-0.969256f * v.x + 1.875991f * v.y) + 0.041556f * v.z,
                (0.055648f * v.x + -0.204043f * v.y) + 1.057311f * v.z);
}

// x, y, z of the CIE 1931 standard observer at 380 + 5 i nm, i = 0..80; the
// table is a switch so that no memory table is indexed at run time
RM_HD vec3 cie_cmf_5nm(int i) {
    switch (i) {
    case 0: return vec3(0.0014f, 0.0000f, 0.0065f);
    case 1: return vec3(0.0022f, 0.0001f, 0.0105f);
    case 2: return vec3(0.0042f, 0.0001f, 0.0201f);
    case 3: return vec3(0.0076f, 0.0002f, 0.0362f);
    case 4: return vec3(0.0143f, 0.0004f, 0.0679f);
    case 5: return vec3(0.0232f, 0.0006f, 0.1102f);
    case 6: return vec3(0.0435f, 0.0012f, 0.2074f);
    case 7: return vec3(0.0776f, 0.0022f, 0.3713f);
    case 8: return vec3(0.1344f, 0.0040f, 0.6456f);
    case 9: return vec3(0.2148f, 0.0073f, 1.0391f);
    case 10: return vec3(0.2839f, 0.0116f, 1.3856f);
    case 11: return vec3(0.3285f, 0.0168f, 1.6230f);
    case 12: return vec3(0.3483f, 0.0230f, 1.7471f);
    case 13: return vec3(0.3481f, 0.0298f, 1.7826f);
    case 14: return vec3(0.3362f, 0.0380f, 1.7721f);
    case 15: return vec3(0.3187f, 0.0480f, 1.7441f);
    case 16: return vec3(0.2908f, 0.0600f, 1.6692f);
    case 17: return vec3(0.2511f, 0.0739f, 1.5281f);
    case 18: return vec3(0.1954f, 0.0910f, 1.2876f);
    case 19: return vec3(0.1421f, 0.1126f, 1.0419f);
    case 20: return vec3(0.0956f, 0.1390f, 0.8130f);
    case 21: return vec3(0.0580f, 0.1693f, 0.6162f);
    case 22: return vec3(0.0320f, 0.2080f, 0.4652f);
    case 23: return vec3(0.0147f, 0.2586f, 0.3533f);
    case 24: return vec3(0.0049f, 0.3230f, 0.2720f);
    case 25: return vec3(0.0024f, 0.4073f, 0.2123f);
    case 26: return vec3(0.0093f, 0.5030f, 0.1582f);
    case 27: return vec3(0.0291f, 0.6082f, 0.1117f);
    case 28: return vec3(0.0633f, 0.7100f, 0.0782f);
    case 29: return vec3(0.1096f, 0.7932f, 0.0573f);
    case 30: return vec3(0.1655f, 0.8620f, 0.0422f);
    case 31: return vec3(0.2257f, 0.9149f, 0.0298f);
    case 32: return vec3(0.2904f, 0.9540f, 0.0203f);
    case 33: return vec3(0.3597f, 0.9803f, 0.0134f);
    case 34: return vec3(0.4334f, 0.9950f, 0.0087f);
    case 35: return vec3(0.5121f, 1.0000f, 0.0057f);
    case 36: return vec3(0.5945f, 0.9950f, 0.0039f);
    case 37: return vec3(0.6784f, 0.9786f, 0.0027f);
    case 38: return vec3(0.7621f, 0.9520f, 0.0021f);
    case 39: return vec3(0.8425f, 0.9154f, 0.0018f);
    case 40: return vec3(0.9163f, 0.8700f, 0.0017f);
    case 41: return vec3(0.9786f, 0.8163f, 0.0014f);
    case 42: return vec3(1.0263f, 0.7570f, 0.0011f);
    case 43: return vec3(1.0567f, 0.6949f, 0.0010f);
    case 44: return vec3(1.0622f, 0.6310f, 0.0008f);
    case 45: return vec3(1.0456f, 0.5668f, 0.0006f);
    case 46: return vec3(1.0026f, 0.5030f, 0.0003f);
    case 47: return vec3(0.9384f, 0.4412f, 0.0002f);
    case 48: return vec3(0.8544f, 0.3810f, 0.0002f);
    case 49: return vec3(0.7514f, 0.3210f, 0.0001f);
    case 50: return vec3(0.6424f, 0.2650f, 0.0000f);
    case 51: return vec3(0.5419f, 0.2170f, 0.0000f);
    case 52: return vec3(0.4479f, 0.1750f, 0.0000f);
    case 53: return vec3(0.3608f, 0.1382f, 0.0000f);
    case 54: return vec3(0.2835f, 0.1070f, 0.0000f);
    case 55: return vec3(0.2187f, 0.0816f, 0.0000f);
    case 56: return vec3(0.1649f, 0.0610f, 0.0000f);
    case 57: return vec3(0.1212f, 0.0446f, 0.0000f);
    case 58: return vec3(0.0874f, 0.0320f, 0.0000f);
    case 59: return vec3(0.0636f, 0.0232f, 0.0000f);
    case 60: return vec3(0.0468f, 0.0170f, 0.0000f);
    case 61: return vec3(0.0329f, 0.0119f, 0.0000f);
    case 62: return vec3(0.0227f, 0.0082f, 0.0000f);
    case 63: return vec3(0.0158f, 0.0057f, 0.0000f);
    case 64: return vec3(0.0114f, 0.0041f, 0.0000f);
    case 65: return vec3(0.0081f, 0.0029f, 0.0000f);
    case 66: return vec3(0.0058f, 0.0021f, 0.0000f);
    case 67: return vec3(0.0041f, 0.0015f, 0.0000f);
    case 68: return vec3(0.0029f, 0.0010f, 0.0000f);
    case 69: return vec3(0.0020f, 0.0007f, 0.0000f);
    case 70: return vec3(0.0014f, 0.0005f, 0.0000f);
    case 71: return vec3(0.0010f, 0.0004f, 0.0000f);
    case 72: return vec3(0.0007f, 0.0002f, 0.0000f);
    case 73: return vec3(0.0005f, 0.0002f, 0.0000f);
    case 74: return vec3(0.0003f, 0.0001f, 0.0000f);
    case 75: return vec3(0.0002f, 0.0001f, 0.0000f);
    case 76: return vec3(0.0002f, 0.0001f, 0.0000f);
    case 77: return vec3(0.0001f, 0.0000f, 0.0000f);
    case 78: return vec3(0.0001f, 0.0000f, 0.0000f);
    case 79: return vec3(0.0001f, 0.0000f, 0.0000f);
    case 80: return vec3(0.0000f, 0.0000f, 0.0000f);
    default: return vec3(0.0f, 0.0f, 0.0f);
    }
}

// linear interpolation of the 5-nm table (color.hpp:131-138); zero from
// 780 nm, and, as there, the truncation toward zero extends the first
// interval slightly below 380 nm
RM_HD vec3 wavelength_to_xyz(float nm) {
    const float ii = (nm - 380.0f) * (1.0f / 5.0f);
    const int i = (int)ii;
    if (i < 0 || i >= 80) return vec3(0.0f);
    const float f = ii - (float)i;
    return (1.0f - f) * cie_cmf_5nm(i) + f * cie_cmf_5nm(i + 1);
}

// 1 / integral over 380-780 nm of the colour matching functions mapped to
// linear sRGB, per channel (1 / nm): a uniform wavelength in that range
// weighted by xyz_to_rgb(cmf) * kCieNormRgb * 400 nm has expectation 1 in
// each channel
constexpr float kCieNormRgbR = 0.00776762f, kCieNormRgbG = 0.00986860f, kCieNormRgbB = 0.01029787f;

}  // namespace rm
