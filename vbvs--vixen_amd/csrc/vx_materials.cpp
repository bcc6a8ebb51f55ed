// vx_materials.cpp -- VIXEN's material-id colours: the uncompressed shader's
// getMaterialColor (shaders/Materials.glsl:18-62) and the rebuild's CPU table
// MaterialIdToColor (libraries/SVO/src/SVORebuild.cpp:68-120). The two agree
// except for ids 33-39, which the shader interpolates and the CPU table lists.
// FP32 throughout, evaluated in the order the sources write it (built with
// -ffp-contract=off, so no step is fused).
#include <cstdint>

#include "vixen_build.h"
#include "vixen_hip.h"

namespace {

struct Named {
    uint32_t id;
    float r, g, b;
};

// The ids both sources give a constant colour (Cornell box 1-20, caves 30-32
// and 40, city 50-61).
constexpr Named kShared[] = {
    {1, 1.0f, 0.0f, 0.0f},    {2, 0.0f, 1.0f, 0.0f},     {3, 0.9f, 0.9f, 0.9f},     {4, 0.85f, 0.85f, 0.85f},
    {5, 0.95f, 0.95f, 0.95f}, {6, 0.8f, 0.8f, 0.8f},     {7, 0.4f, 0.4f, 0.4f},     {10, 0.8f, 0.6f, 0.2f},
    {11, 0.6f, 0.8f, 0.9f},   {20, 1.0f, 0.98f, 0.9f},   {30, 0.5f, 0.45f, 0.4f},   {31, 0.6f, 0.55f, 0.5f},
    {32, 0.55f, 0.5f, 0.45f}, {40, 0.8f, 0.6f, 0.2f},    {50, 0.2f, 0.2f, 0.22f},   {60, 0.6f, 0.58f, 0.55f},
    {61, 0.7f, 0.85f, 0.95f},
};

// The CPU table's stone variants (SVORebuild.cpp:86-92).
constexpr Named kCpuStone[] = {
    {33, 0.52f, 0.48f, 0.42f}, {34, 0.48f, 0.44f, 0.38f}, {35, 0.45f, 0.42f, 0.36f}, {36, 0.58f, 0.52f, 0.46f},
    {37, 0.62f, 0.56f, 0.48f}, {38, 0.65f, 0.58f, 0.5f},  {39, 0.68f, 0.6f, 0.52f},
};

bool lookup(const Named* t, int n, uint32_t id, float rgb[3])
{
    for (int i = 0; i < n; ++i)
        if (t[i].id == id) {
            rgb[0] = t[i].r;
            rgb[1] = t[i].g;
            rgb[2] = t[i].b;
            return true;
        }
    return false;
}

void set(float rgb[3], float r, float g, float b)
{
    rgb[0] = r;
    rgb[1] = g;
    rgb[2] = b;
}

// The fallback both sources share: hue (id mod 256) / 256 on an HSV wheel of
// saturation 0.7 and value 0.8.
void hsv_wheel(uint32_t id, float rgb[3])
{
    const float hue = static_cast<float>(id % 256u) / 256.0f;
    const float h = hue * 6.0f;
    const int i = static_cast<int>(h);
    const float f = h - static_cast<float>(i);
    const float p = 0.8f * 0.3f;
    const float q = 0.8f * (1.0f - 0.7f * f);
    const float t = 0.8f * (1.0f - 0.7f * (1.0f - f));
    switch (i) {
    case 0: set(rgb, 0.8f, t, p); break;
    case 1: set(rgb, q, 0.8f, p); break;
    case 2: set(rgb, p, 0.8f, t); break;
    case 3: set(rgb, p, q, 0.8f); break;
    case 4: set(rgb, t, p, 0.8f); break;
    default: set(rgb, 0.8f, p, q); break;
    }
}

} // namespace

extern "C" {

void vx_material_color(uint32_t id, float rgb[3])
{
    if (lookup(kShared, sizeof kShared / sizeof kShared[0], id, rgb)) return;
    if (id >= 33u && id <= 39u) {
        // mix(vec3(0.52, 0.48, 0.42), vec3(0.68, 0.6, 0.52), t), mix(x, y, a) = x (1 - a) + y a
        const float t = static_cast<float>(id - 33u) / 6.0f;
        const float a[3] = {0.52f, 0.48f, 0.42f}, b[3] = {0.68f, 0.6f, 0.52f};
        for (int k = 0; k < 3; ++k) rgb[k] = a[k] * (1.0f - t) + b[k] * t;
        return;
    }
    hsv_wheel(id, rgb);
}

void vx_material_palette(float palette[256 * 3])
{
    for (uint32_t id = 0; id < 256u; ++id) vx_material_color(id, palette + 3 * id);
}

void vx_material_id_to_color(uint32_t id, float rgb[3])
{
    if (lookup(kShared, sizeof kShared / sizeof kShared[0], id, rgb)) return;
    if (lookup(kCpuStone, sizeof kCpuStone / sizeof kCpuStone[0], id, rgb)) return;
    hsv_wheel(id, rgb);
}

} // extern "C"
