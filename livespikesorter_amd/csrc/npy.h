// Minimal .npy (format 1.0-3.0) reader for the artifact directory.  Accepts little-endian
// float32/float64/int16/int32/int64/uint8 arrays and converts them to float32 in C order;
// Fortran-ordered files are transposed so that the result is the array numpy would return.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace lss {

struct NpyArray {
    std::vector<int64_t> shape;
    std::vector<float> data;
};

bool npy_load_f32(const std::string& path, NpyArray& out, std::string& err);

}  // namespace lss
