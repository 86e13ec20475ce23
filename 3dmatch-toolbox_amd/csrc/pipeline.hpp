// Internal helpers shared by the registration server and the entry points it
// drives: the host half of the grid-size rule and the scratch sizes of the
// matching and RANSAC entry points (so a server can reserve them up front).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace tdm {

// dims = round(extent / vs) + 1 + 2 pad, origin = min - pad vs + vs/2 in
// float arithmetic (core/demo.cu:116-122), from box = {min xyz, max xyz}.
int grid_from_box(const float box[6], float voxel_size, int padding, int dims[3],
                  float origin[3]);

// Scratch (context slot 2) used by tdm_match_nn for num_a x num_b of width dim.
size_t match_scratch_bytes(int num_a, int num_b, int dim);
// Scratch (context slot 3) used by tdm_ransac_rt for max_trials.
size_t ransac_scratch_bytes(long max_trials);

// Decodes a 16-bit grayscale PNG (depth-fusion/utils.hpp:74-86's reader).
int read_png_u16(const char *path, std::vector<uint16_t> &pix, int &H, int &W);

}  // namespace tdm
