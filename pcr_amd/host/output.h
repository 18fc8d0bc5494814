// output.h -- tone mapping and PNG output (see output.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "scene.h"

namespace pcr {

void tonemap_reinhard(Vec3 &c);
void tonemap_aces(Vec3 &c);
uint8_t quantize_clamped(float v);
// linear frame -> tone-mapped, clamped, quantised RGB8
std::vector<uint8_t> frame_to_rgb8(const std::vector<Vec3> &frame, bool aces);
bool encode_png(const std::vector<uint8_t> &rgb, int w, int h,
                const std::vector<std::pair<std::string, std::string>> &text, std::vector<uint8_t> &png,
                std::string *err);
bool write_file(const std::string &path, const std::vector<uint8_t> &data, std::string *err);

}  // namespace pcr
