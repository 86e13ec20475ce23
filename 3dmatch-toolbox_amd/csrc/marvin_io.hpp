// .marvin tensor files (core/marvin.hpp:2581-2611 read, 2764-2798 write):
// a sequence of {u8 type id, u32 sizeof(element), i32 name length, name
// bytes (no NUL), i32 ndims, i32 dims[ndims], raw little-endian data}.
// Type ids: 0 half, 1 float, 2 double (others are skipped when read).
#pragma once

#include <string>
#include <vector>

namespace tdm {

struct MarvinTensor {
    std::string name;
    std::vector<int> dims;
    std::vector<float> data;  // converted to float
};

int marvin_read(const char *path, std::vector<MarvinTensor> &out);
int marvin_write(const char *path, const std::vector<MarvinTensor> &ts);
float half_to_float(uint16_t h);

}  // namespace tdm
