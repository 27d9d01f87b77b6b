// Minimal .npy (format 1.0-3.0) reader for the artifact directory.  Accepts little-endian
// float32/float64/int16/int32/int64/uint8/bool arrays and converts them to float32 (npy_load_f32)
// or int64 (npy_load_i64) in C order; Fortran-ordered files are transposed so that the result is
// the array numpy would return.  Malformed files (bad header, negative or overflowing shape, data
// shorter than the header says) are reported through err; nothing throws.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace lss {

struct NpyArray {
    std::vector<int64_t> shape;
    std::vector<float> data;
};

struct NpyArrayI64 {
    std::vector<int64_t> shape;
    std::vector<int64_t> data;
};

bool npy_load_f32(const std::string& path, NpyArray& out, std::string& err);
// Integer arrays (index tables such as iCC.npy / iU.npy); float files must hold integral values.
bool npy_load_i64(const std::string& path, NpyArrayI64& out, std::string& err);

}  // namespace lss
