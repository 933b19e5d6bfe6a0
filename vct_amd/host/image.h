// image.h -- a decoded diffuse map: RGBA8 levels back to back (vct_texture's
// layout), shared by the DDS and PNG loaders.
#pragma once
#include <cstdint>
#include <vector>

struct vct_image {
    int width = 0, height = 0, levels = 0;
    std::vector<uint32_t> texels;
};
